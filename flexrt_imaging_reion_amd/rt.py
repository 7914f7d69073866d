"""Host-side mirror of the reference's per-step RT routines on the GPU.

``FlexRT`` owns one ``frt_ctx`` of libflexrt_gpu.so.  Its methods carry the
names of the reference routines they replace (rt_funcs.cc, ray_tracing.cc,
gas_funcs.cc) and take and return numpy arrays in the reference's layouts.
Every call goes through the C ABI; a missing library raises at construction.
"""
import ctypes as C
import math
import os

import numpy as np

from . import _lib
from ._abi import (EXTRA_FLAG_FIELDS, EXTRA_FLOAT_FIELDS, FIELD_ID, PER_BIN, PREC_FP64, PREC_MIXED, FrtOtf,
                   FrtConfig, FrtStepStats, config_struct, make_extras, make_ray_view, make_subgrid,
                   make_subgrid_state)


class FlexRTError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("frt error %d: %s" % (code, msg))
        self.code = code


def default_config(n):
    """The reference's compile-time configuration (user_inputs.h) for an n^3 grid."""
    lib = _lib.load()
    s = FrtConfig()
    rc = lib.frt_config_default(C.byref(s), n)
    if rc:
        raise FlexRTError(rc, "frt_config_default")
    return {name: getattr(s, name) for name, _ in FrtConfig._fields_}


def checkpoint_paths(output_dir, zz):
    """File names of the driver's checkpoint at redshift zz
    (gas_funcs.cc:617-645): output_dir (concatenated as given, so it carries
    its own separator) + "gas_z=" / "ray_z=" / "tre_z=" + zz with six decimals
    less the last two, zero-padded to two integer digits below z = 10."""
    zs = "%f" % float(zz)
    pad = "0" if float(zz) < 10.0 else ""
    return {k: str(output_dir) + k + "_z=" + pad + zs[:-2] for k in ("gas", "ray", "tre")}


def gadget_path(output_dir, current_time, step_time, hh, omega_m):
    """write_gadget_binary's file name (io_funcs.cc:465-467): the redshift of
    current_time - step_time/2 by toRedshift (cosmo_funcs.cc:108-113, an
    Einstein-de Sitter + Lambda closed form), six decimals."""
    H0 = 3.24086e-18
    t = float(current_time) - float(step_time) / 2.0
    atime = (math.sqrt(omega_m / (1.0 - omega_m)) * math.sinh(1.5 * H0 * hh * math.sqrt(1.0 - omega_m) * t)) ** (2.0 / 3.0)
    return str(output_dir) + "small_gas_rt_z=" + "%f" % (1.0 / atime - 1.0)


def vif_path(output_dir, zz, z_rt):
    """write_vIF_binary's file name (io_funcs.cc:399-414): the digits are those
    of z_rt = 1/a_rt - 1, the zero padding follows the snapshot redshift zz."""
    pad = "0" if float(zz) < 10.0 else ""
    return str(output_dir) + "vIF_z=" + pad + ("%f" % float(z_rt))[:-2]


class Checkpoint:
    """A device checkpoint (frt_checkpoint_save); freed with free() or when
    collected."""

    def __init__(self, lib, handle):
        self._lib = lib
        self.handle = handle

    def free(self):
        if self.handle:
            self._lib.frt_checkpoint_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class FlexRT:
    """One GPU RT context (device state resident in HBM)."""

    def __init__(self, cfg):
        self._lib = _lib.load()
        self.cfg = dict(cfg)
        s = config_struct(self.cfg)
        h = C.c_void_p()
        rc = self._lib.frt_create(C.byref(s), C.byref(h))
        if rc:
            raise FlexRTError(rc, "frt_create failed (see stderr)")
        self._h = h
        self.n = int(self.cfg["n"])
        self.nfreq = int(self.cfg["nfreq"])
        self.path_length = int(self.cfg["path_length"])

    # -- plumbing ------------------------------------------------------------
    def _check(self, rc, what):
        if rc:
            msg = self._lib.frt_last_error(self._h)
            raise FlexRTError(rc, "%s: %s" % (what, msg.decode() if msg else ""))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.frt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- inputs ----------------------------------------------------------------
    def set_spectrum(self, nu_phot, spectrum):
        nu = np.ascontiguousarray(nu_phot, dtype=np.float64)
        sp = np.ascontiguousarray(spectrum, dtype=np.float32)
        assert nu.size == self.nfreq and sp.size == self.nfreq
        self._check(self._lib.frt_set_spectrum(
            self._h, nu.ctypes.data_as(C.POINTER(C.c_double)),
            sp.ctypes.data_as(C.POINTER(C.c_float))), "set_spectrum")

    def set_sources(self, cells, lum):
        c = np.ascontiguousarray(cells, dtype=np.int32)
        lu = np.ascontiguousarray(lum, dtype=np.float64)
        assert c.size == lu.size
        self._srcs = c.copy()
        self._check(self._lib.frt_set_sources(
            self._h, int(c.size), c.ctypes.data_as(C.POINTER(C.c_int32)),
            lu.ctypes.data_as(C.POINTER(C.c_double))), "set_sources")

    def _field_size(self, name):
        n3 = self.n ** 3
        return n3 * self.nfreq if name in PER_BIN else n3

    def upload_field(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32).reshape(-1)
        assert a.size == self._field_size(name), name
        self._check(self._lib.frt_upload_field(
            self._h, FIELD_ID[name], a.ctypes.data_as(C.POINTER(C.c_float))), "upload " + name)

    def download_field(self, name):
        a = np.empty(self._field_size(name), dtype=np.float32)
        self._check(self._lib.frt_download_field(
            self._h, FIELD_ID[name], a.ctypes.data_as(C.POINTER(C.c_float))), "download " + name)
        return a.reshape(-1, self.nfreq) if name in PER_BIN else a

    def set_time(self, t):
        self._check(self._lib.frt_set_time(self._h, float(t)), "set_time")

    # ---- sub-grid I-front model (cfg["subgrid"] = 1) -------------------------------
    def set_subgrid(self, sg):
        """RadHydro tables and constants (a dict as subgrid.make_params returns)."""
        st, keep = make_subgrid(sg)
        self._check(self._lib.frt_set_subgrid(self._h, C.byref(st)), "set_subgrid")

    def set_subgrid_delta_nl(self, d):
        a = np.ascontiguousarray(d, dtype=np.float32)
        self._check(self._lib.frt_set_subgrid_delta_nl(self._h, a.ctypes.data_as(C.POINTER(C.c_float))),
                    "set_subgrid_delta_nl")

    def download_subgrid_state(self):
        """The per-cell I-front state, row-major (tcross_new, tcross_check and
        fh1_nodes with 10, 10 and 5 values per cell)."""
        st, arrays = make_subgrid_state(self.n ** 3)
        self._check(self._lib.frt_download_subgrid_state(self._h, C.byref(st)), "download_subgrid_state")
        return arrays

    def upload_subgrid_state(self, arrays):
        st, keep = make_subgrid_state(self.n ** 3, arrays)
        self._check(self._lib.frt_upload_subgrid_state(self._h, C.byref(st)), "upload_subgrid_state")

    def set_step(self, step):
        """Restart: completed steps (they key the casting and merging angle streams)."""
        self._check(self._lib.frt_set_step(self._h, int(step)), "set_step")

    def set_opacity_dither(self, seed):
        """Rounding-sensitivity measurement: one-ulp dither of the stored
        opacities after every C2Ray pass (seed 0 turns it off)."""
        self._check(self._lib.frt_set_opacity_dither(self._h, int(seed)), "set_opacity_dither")

    def set_dt(self, dt):
        """update_dt (rt_funcs.cc:9-12): the driver's new time step."""
        self._check(self._lib.frt_set_dt(self._h, float(dt)), "set_dt")

    def cell_size(self):
        """dx[0], cm (grows with update_cosmo_expansion)."""
        v = C.c_float()
        self._check(self._lib.frt_get_cell_size(self._h, C.byref(v)), "get_cell_size")
        return v.value

    def update_cosmo_expansion(self, a_current, a_prev):
        """update_cosmo_expansion (gas_funcs.cc:460-552) for the driver's scale
        factors now and one dt earlier; needs rho uploaded."""
        self._check(self._lib.frt_update_cosmo_expansion(self._h, float(a_current), float(a_prev)),
                    "update_cosmo_expansion")

    def update_hydro_step(self, zz, a_rt, avg_rho, dt_cosmic):
        """Device part of update_hydro_step (gas_funcs.cc:554-752): call after
        uploading the next snapshot's comoving density as "rho" (and its
        sources with set_sources).  avg_rho is the driver's volume-average
        density after its a_rt^-3 scaling, dt_cosmic = cosmic_time(aa_next) -
        cosmic_time(aa); they set ddelta_dt and rho_prev (:700-738)."""
        self._check(self._lib.frt_update_hydro_step(self._h, float(zz), float(a_rt), float(avg_rho),
                                                    float(dt_cosmic)), "update_hydro_step")

    def otf_averages(self, avg_rho, temp_0=100.0):
        """write_otf's global averages (io_funcs.cc:500-632), reduced on the GPU."""
        o = FrtOtf()
        self._check(self._lib.frt_otf_averages(self._h, float(avg_rho), float(temp_0), C.byref(o)), "otf_averages")
        return o.as_dict()

    def write_otf(self, path, avg_rho, temp_0=100.0):
        """write_otf: appends this step's block to the text file `path`."""
        self._check(self._lib.frt_write_otf(self._h, os.fsencode(path), float(avg_rho), float(temp_0)), "write_otf")

    def get_time(self):
        t = C.c_float()
        self._check(self._lib.frt_get_time(self._h, C.byref(t)), "get_time")
        return t.value

    # -- reference routines ----------------------------------------------------------
    def init_rt(self):
        self._check(self._lib.frt_init_rt(self._h), "init_rt")

    def get_nphot_min_max(self):
        a, b = C.c_double(), C.c_double()
        self._check(self._lib.frt_get_nphot_min_max(self._h, C.byref(a), C.byref(b)), "get_nphot_min_max")
        return a.value, b.value

    def new_rays(self):
        self._check(self._lib.frt_new_rays(self._h), "new_rays")

    def update_rays(self):
        self._check(self._lib.frt_update_rays(self._h), "update_rays")

    def find_avg_unu(self):
        self._check(self._lib.frt_find_avg_unu(self._h), "find_avg_unu")

    def merge_rays(self):
        self._check(self._lib.frt_merge_rays(self._h), "merge_rays")

    def split_rays(self):
        self._check(self._lib.frt_split_rays(self._h), "split_rays")

    def update_thermal(self):
        self._check(self._lib.frt_update_thermal(self._h), "update_thermal")

    def step(self):
        st = FrtStepStats()
        self._check(self._lib.frt_step(self._h, C.byref(st)), "step")
        return st.as_dict()

    def stats(self):
        st = FrtStepStats()
        self._check(self._lib.frt_get_stats(self._h, C.byref(st)), "get_stats")
        return st.as_dict()

    # -- step graph and launch tuning ---------------------------------------------------
    def set_step_graph(self, on=True):
        """Capture frt_step into a HIP graph on first use and replay it."""
        self._check(self._lib.frt_set_step_graph(self._h, 1 if on else 0), "set_step_graph")

    TUNING_KEYS = ("deposit_queue", "march_queue", "queue_chunk", "blocks_per_cu", "grid_cap", "fp64_chunk")

    def set_tuning(self, **kw):
        for k, v in kw.items():
            self._check(self._lib.frt_set_tuning(self._h, k.encode(), int(v)), "set_tuning(%s)" % k)

    def tuning(self):
        out = {}
        for k in self.TUNING_KEYS + ("num_cu",):
            v = C.c_int64()
            self._check(self._lib.frt_get_tuning(self._h, k.encode(), C.byref(v)), "get_tuning(%s)" % k)
            out[k] = v.value
        return out

    def selftest_scan_sort(self, keys=None, bits=32, counts=None):
        """The library's device-count scan and stable sort on host arrays:
        (sorted keys, permutation, exclusive scan, total)."""
        n = len(keys) if keys is not None else len(counts)
        P32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        PU = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
        k = ko = perm = c = so = None
        if keys is not None:
            k = np.ascontiguousarray(keys, dtype=np.uint32)
            ko = np.zeros(n, dtype=np.uint32)
            perm = np.zeros(n, dtype=np.int32)
        if counts is not None:
            c = np.ascontiguousarray(counts, dtype=np.int32)
            so = np.zeros(n, dtype=np.int32)
        tot = C.c_int64(-1)
        self._check(self._lib.frt_selftest_scan_sort(
            self._h, n, int(bits), PU(k) if k is not None else None, PU(ko) if ko is not None else None,
            P32(perm) if perm is not None else None, P32(c) if c is not None else None,
            P32(so) if so is not None else None, C.byref(tot)), "selftest_scan_sort")
        return ko, perm, so, tot.value

    def selftest_sort_cells(self, cells):
        """The step's cell sort of segments on a host array of cell indices:
        (slot of every item, start and end of every cell's slot range)."""
        c = np.ascontiguousarray(cells, dtype=np.uint32)
        n3 = self.n ** 3
        slot = np.zeros(len(c), dtype=np.int32)
        cs = np.zeros(n3, dtype=np.int32)
        ce = np.zeros(n3, dtype=np.int32)
        P32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        self._check(self._lib.frt_selftest_sort_cells(self._h, len(c), c.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                      P32(slot), P32(cs), P32(ce)), "selftest_sort_cells")
        return slot, cs, ce

    # -- rays --------------------------------------------------------------------------
    def ray_count(self):
        n = C.c_int64()
        self._check(self._lib.frt_ray_count(self._h, C.byref(n)), "ray_count")
        return n.value

    def rays(self):
        n = self.ray_count()
        v, arrays = make_ray_view(n, self.nfreq)
        self._check(self._lib.frt_download_rays(self._h, C.byref(v)), "download_rays")
        return {k: a[: v.count] for k, a in arrays.items()}

    def upload_rays(self, rays):
        n = len(rays["cell"])
        v, arrays = make_ray_view(n, self.nfreq, rays)
        self._keep = arrays
        self._check(self._lib.frt_upload_rays(self._h, C.byref(v)), "upload_rays")

    def paths(self):
        n = self.ray_count()
        P = self.path_length
        end = np.zeros(n, dtype=np.int32)
        cells = np.zeros((n, P), dtype=np.int32)
        dists = np.zeros((n, P), dtype=np.float32)
        self._check(self._lib.frt_download_paths(
            self._h, n, end.ctypes.data_as(C.POINTER(C.c_int32)),
            cells.ctypes.data_as(C.POINTER(C.c_int32)),
            dists.ctypes.data_as(C.POINTER(C.c_float))), "download_paths")
        return end, cells, dists

    # -- restart files (io_funcs.cc formats) --------------------------------------------
    def write_ray_file(self, path):
        """write_rays_binary (io_funcs.cc:274-303)."""
        self._check(self._lib.frt_write_ray_file(self._h, os.fsencode(path)), "write_ray_file")

    def read_ray_file(self, path):
        """read_rays_binary (io_funcs.cc:119-145): replaces the ray buffer."""
        self._check(self._lib.frt_read_ray_file(self._h, os.fsencode(path)), "read_ray_file")

    def write_gas_file(self, path, extras=None):
        """write_gas_binary (io_funcs.cc:358-393); extras: the fields the
        library does not evolve (see frt_file_extras), row-major."""
        ex, keep = make_extras(self.n ** 3, extras)
        self._check(self._lib.frt_write_gas_file(self._h, os.fsencode(path), C.byref(ex)), "write_gas_file")

    def read_gas_file(self, path):
        """read_grid_binary (io_funcs.cc:81-116); returns the extras of the file."""
        ex, arrays = make_extras(self.n ** 3, alloc=True)
        self._check(self._lib.frt_read_gas_file(self._h, os.fsencode(path), C.byref(ex)), "read_gas_file")
        return {k: arrays[k] for k in EXTRA_FLOAT_FIELDS}

    def write_treion_file(self, path, extras=None):
        """write_treion (io_funcs.cc:338-355)."""
        ex, keep = make_extras(self.n ** 3, extras)
        self._check(self._lib.frt_write_treion_file(self._h, os.fsencode(path), C.byref(ex)),
                    "write_treion_file")

    def read_treion_file(self, path):
        """read_treion (io_funcs.cc:65-77); returns tcross_flag, neutral_flag."""
        ex, arrays = make_extras(self.n ** 3, alloc=True)
        self._check(self._lib.frt_read_treion_file(self._h, os.fsencode(path), C.byref(ex)),
                    "read_treion_file")
        return {k: arrays[k] for k in EXTRA_FLAG_FIELDS}

    def write_vif_file(self, path):
        """write_vIF_binary (io_funcs.cc:396-430)."""
        self._check(self._lib.frt_write_vif_file(self._h, os.fsencode(path)), "write_vif_file")

    # -- device checkpoints -----------------------------------------------------------------
    def checkpoint(self):
        """The whole state a later step depends on, kept on the device."""
        h = C.c_void_p()
        self._check(self._lib.frt_checkpoint_save(self._h, C.byref(h)), "checkpoint_save")
        return Checkpoint(self._lib, h)

    def restore(self, ck):
        self._check(self._lib.frt_checkpoint_restore(self._h, ck.handle), "checkpoint_restore")

    # -- adjoint of find_avg_unu ------------------------------------------------
    def adj_enable(self, on=True):
        """Record an adjoint tape in every following find_avg_unu."""
        self._check(self._lib.frt_adj_enable(self._h, 1 if on else 0), "adj_enable")

    def adj_find_avg_unu(self, a_nh1):
        """Gradient of <a_nh1, n_HI after the last find_avg_unu> with respect to
        the rays' starting photons ([ray][nfreq], stored units), clump, nH, nHe
        and the starting nH1, nHe1 (row-major cells); paths, dumps and
        recombination coefficients held fixed."""
        n3 = self.n ** 3
        a = np.ascontiguousarray(np.asarray(a_nh1, dtype=np.float64).reshape(n3))
        out = {k: np.zeros(n3) for k in ("clump", "nH", "nHe", "nH1", "nHe1")}
        nr = C.c_int64()
        self._check(self._lib.frt_adj_ray_count(self._h, C.byref(nr)), "adj_ray_count")
        nph = np.zeros(nr.value * self.nfreq + 1)
        P = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
        self._check(self._lib.frt_adj_find_avg_unu(self._h, P(a), P(nph), P(out["clump"]), P(out["nH"]),
                                                   P(out["nHe"]), P(out["nH1"]), P(out["nHe1"])),
                    "adj_find_avg_unu")
        out["nphot"] = nph[: nr.value * self.nfreq].reshape(nr.value, self.nfreq)
        return out

    # -- adjoint across steps ---------------------------------------------------------
    def adj_run_begin(self):
        """Start a reverse sweep from the current state (needs adj_enable)."""
        self._check(self._lib.frt_adj_run_begin(self._h), "adj_run_begin")

    def adj_seed(self, a_nH1=None, a_nH=None):
        """Add the adjoint of the current n_HI and nH (row-major cells)."""
        n3 = self.n ** 3
        P = lambda x: None if x is None else np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(n3))
        a, b = P(a_nH1), P(a_nH)
        ptr = lambda x: x.ctypes.data_as(C.POINTER(C.c_double)) if x is not None else None
        self._check(self._lib.frt_adj_seed(self._h, ptr(a), ptr(b)), "adj_seed")

    def adj_step_reverse(self):
        """Reverse the step just run with the tape on."""
        self._check(self._lib.frt_adj_step_reverse(self._h), "adj_step_reverse")

    def adj_run_finish(self, init_ray_dt):
        """Parameter adjoints once every step is reversed and the context is
        back in its initial state: d/dL per source, d/dC and d/dlog rho per
        row-major cell."""
        n3 = self.n ** 3
        nsrc = len(self._srcs) if hasattr(self, "_srcs") else 0
        gl = np.zeros(max(nsrc, 1))
        gc = np.zeros(n3)
        gr = np.zeros(n3)
        P = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
        self._check(self._lib.frt_adj_run_finish(self._h, float(init_ray_dt), P(gl), P(gc), P(gr)),
                    "adj_run_finish")
        return {"lum": gl[:nsrc], "clump": gc, "log_rho": gr}

    # -- MP-Gadget coupling (io_funcs.cc:434-496) ----------------------------------------
    def update_gadget_var(self):
        """Adds this step's Gamma_HI dt and heat_rate dt / rho (needs rho)."""
        self._check(self._lib.frt_update_gadget_var(self._h), "update_gadget_var")

    def write_gadget_file(self, path):
        """write_gadget_binary: averages over the step time, in place, then the
        two row-major cubes (gadget_gamma, gadget_heat)."""
        self._check(self._lib.frt_write_gadget_file(self._h, os.fsencode(path)), "write_gadget_file")

    def reset_gadget_var(self):
        self._check(self._lib.frt_reset_gadget_var(self._h), "reset_gadget_var")

    def gadget_step_time(self):
        v = C.c_float()
        self._check(self._lib.frt_get_gadget_step_time(self._h, C.byref(v)), "get_gadget_step_time")
        return v.value

    def write_checkpoint(self, output_dir, zz, extras=None, rays=True):
        """The end-of-hydro-step output of the driver (gas_funcs.cc:617-656):
        gas file, ray file (the driver writes it only for some snapshots, hence
        `rays`), treion file, named as checkpoint_paths names them.  Returns
        the paths written."""
        paths = checkpoint_paths(output_dir, zz)
        self.write_gas_file(paths["gas"], extras)
        if rays:
            self.write_ray_file(paths["ray"])
        self.write_treion_file(paths["tre"], extras)
        return paths

    # -- measurement -------------------------------------------------------------------
    def set_profiling(self, on):
        self._check(self._lib.frt_set_profiling(self._h, 1 if on else 0), "set_profiling")

    def kernel_time(self, name):
        ms, n = C.c_double(), C.c_int64()
        self._check(self._lib.frt_kernel_time(self._h, name.encode(), C.byref(ms), C.byref(n)),
                    "kernel_time")
        return ms.value, n.value

    def reset_kernel_times(self):
        self._check(self._lib.frt_reset_kernel_times(self._h), "reset_kernel_times")

    def deposit_bytes(self):
        b, n = C.c_double(), C.c_int64()
        self._check(self._lib.frt_deposit_bytes(self._h, C.byref(b), C.byref(n)), "deposit_bytes")
        return b.value, n.value

    def copy_kernel(self, nbytes, repeats=1):
        self._check(self._lib.frt_copy_kernel(self._h, int(nbytes), int(repeats)), "copy_kernel")

    def synchronize(self):
        self._check(self._lib.frt_synchronize(self._h), "synchronize")


__all__ = ["FlexRT", "FlexRTError", "default_config", "PREC_FP64", "PREC_MIXED"]
