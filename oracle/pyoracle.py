"""ctypes binding of the CPU restatement (oracle/frt_oracle.cc).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product path.  ``Oracle`` has the
same methods as ``flexrt_imaging_reion_amd.rt.FlexRT`` so one driver runs
either.
"""
import ctypes as C
import os

import numpy as np

from flexrt_imaging_reion_amd._abi import (FIELD_ID, PER_BIN, FrtOtf, FrtRayView, FrtStepStats, FrtSubgrid,
                                           FrtSubgridState, config_struct, make_ray_view, make_subgrid,
                                           make_subgrid_state)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None

_VP = C.c_void_p
_PROTOS = {
    "orc_create": (_VP, [C.c_void_p]),
    "orc_destroy": (None, [_VP]),
    "orc_set_spectrum": (None, [_VP, C.POINTER(C.c_double), C.POINTER(C.c_float)]),
    "orc_set_sources": (None, [_VP, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "orc_upload_field": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_float)]),
    "orc_download_field": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_float)]),
    "orc_set_time": (None, [_VP, C.c_float]),
    "orc_get_time": (C.c_float, [_VP]),
    "orc_init_rt": (None, [_VP]),
    "orc_get_nphot_min_max": (None, [_VP, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "orc_new_rays": (None, [_VP]),
    "orc_step": (C.c_int, [_VP, C.POINTER(FrtStepStats)]),
    "orc_update_rays": (C.c_int, [_VP]),
    "orc_find_avg_unu": (None, [_VP]),
    "orc_merge_rays": (None, [_VP]),
    "orc_split_rays": (None, [_VP]),
    "orc_update_thermal": (None, [_VP]),
    "orc_get_stats": (None, [_VP, C.POINTER(FrtStepStats)]),
    "orc_ray_count": (C.c_int64, [_VP]),
    "orc_download_rays": (C.c_int, [_VP, C.POINTER(FrtRayView)]),
    "orc_upload_rays": (C.c_int, [_VP, C.POINTER(FrtRayView)]),
    "orc_download_paths": (C.c_int, [_VP, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_float)]),
    "orc_dm_exp": (C.c_double, [C.c_double]),
    "orc_cell_index": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_cell_ijk": (None, [C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "orc_tile_key": (C.c_uint, [C.c_int, C.c_int]),
    "orc_fexp_neg": (None, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64]),
    "orc_dm_log": (C.c_double, [C.c_double]),
    "orc_dm_pow": (C.c_double, [C.c_double, C.c_double]),
    "orc_dm_sin": (C.c_double, [C.c_double]),
    "orc_dm_cos": (C.c_double, [C.c_double]),
    "orc_dm_atan2": (C.c_double, [C.c_double, C.c_double]),
    "orc_pix2vec": (None, [C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "orc_vec2pix": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double]),
    "orc_rate": (C.c_double, [C.c_int, C.c_double]),
    "orc_hubble": (C.c_float, [C.c_float, C.c_double, C.c_double, C.c_double]),
    "orc_rotate": (None, [C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "orc_c2ray_cell": (None, [C.c_float] * 9 + [C.c_double, C.c_double, C.c_float, C.POINTER(C.c_float)]),
    "orc_write_ray_file": (C.c_int, [C.c_void_p, C.c_char_p]),
    "orc_update_cosmo_expansion": (None, [C.c_void_p, C.c_float, C.c_float]),
    "orc_set_dt": (None, [C.c_void_p, C.c_float]),
    "orc_update_gadget_var": (None, [C.c_void_p]),
    "orc_gadget_average": (None, [C.c_void_p]),
    "orc_reset_gadget_var": (None, [C.c_void_p]),
    "orc_get_gadget_step_time": (C.c_float, [C.c_void_p]),
    "orc_otf_averages": (None, [C.c_void_p, C.c_float, C.c_float, C.POINTER(FrtOtf)]),
    "orc_set_step": (None, [C.c_void_p, C.c_int64]),
    "orc_set_opacity_dither": (None, [C.c_void_p, C.c_uint64]),
    "orc_set_threads": (C.c_int, [C.c_int]),
    "orc_set_nphot_merge": (None, [C.c_void_p, C.c_float]),
    "orc_update_hydro_step": (None, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float]),
    "orc_get_cell_size": (C.c_float, [C.c_void_p]),
    "orc_itheta_hash": (C.c_uint32, [C.c_uint64, C.c_uint64, C.c_uint64]),
    "orc_derived": (None, [_VP, C.POINTER(C.c_double)]),
    "orc_set_subgrid": (C.c_int, [_VP, C.POINTER(FrtSubgrid)]),
    "orc_set_subgrid_delta_nl": (C.c_int, [_VP, C.POINTER(C.c_float)]),
    "orc_download_subgrid_state": (C.c_int, [_VP, C.POINTER(FrtSubgridState)]),
    "orc_upload_subgrid_state": (C.c_int, [_VP, C.POINTER(FrtSubgridState)]),
    "orc_sg_deltasig": (C.c_float, [C.c_float, C.c_float]),
}

RATES = {"sigmapi_H1": 0, "sigmapi_He1": 1, "alphaB_H2": 2, "alphaB_He2": 3, "alphaA_H2": 4,
         "alphaA_He2": 5, "Dalpha_He2": 6, "rcrB_H2": 7, "rcrB_He2": 8, "rcrA_H2": 9,
         "free_free_rate": 10, "compton_rate": 11, "sigmapi_He2": 12, "alphaB_He3": 13, "alphaA_He3": 14}


_omp_lib = None
OMP_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liboracle_omp.so")


def _bind(path):
    lib = C.CDLL(path)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def load(omp=False):
    """The serial checker, or with omp=True its OpenMP build (the timed CPU
    baseline: the reference's parallel loops, sums in scheduling order)."""
    global _lib, _omp_lib
    if omp:
        if _omp_lib is None:
            if not os.path.exists(OMP_LIB):
                raise FileNotFoundError("oracle/liboracle_omp.so is not built; run __graft_entry__.build()")
            _omp_lib = _bind(OMP_LIB)
        return _omp_lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise FileNotFoundError("oracle/liboracle.so is not built; run __graft_entry__.build()")
        _lib = _bind(LIB)
    return _lib


def set_threads(n):
    """Threads of the OpenMP build; returns the count in use."""
    return load(omp=True).orc_set_threads(int(n))


class Oracle:
    """CPU restatement of the RT step with the FlexRT method names (serial
    unless omp=True)."""

    def __init__(self, cfg, omp=False):
        self._lib = load(omp)
        self.cfg = dict(cfg)
        s = config_struct(self.cfg)
        h = self._lib.orc_create(C.byref(s))
        if not h:
            raise ValueError("oracle rejected the configuration")
        self._h = h
        self.n = int(self.cfg["n"])
        self.nfreq = int(self.cfg["nfreq"])
        self.path_length = int(self.cfg["path_length"])

    def close(self):
        if getattr(self, "_h", None):
            self._lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_spectrum(self, nu_phot, spectrum):
        nu = np.ascontiguousarray(nu_phot, dtype=np.float64)
        sp = np.ascontiguousarray(spectrum, dtype=np.float32)
        self._lib.orc_set_spectrum(self._h, nu.ctypes.data_as(C.POINTER(C.c_double)),
                                   sp.ctypes.data_as(C.POINTER(C.c_float)))

    def set_sources(self, cells, lum):
        c = np.ascontiguousarray(cells, dtype=np.int32)
        lu = np.ascontiguousarray(lum, dtype=np.float64)
        self._lib.orc_set_sources(self._h, int(c.size), c.ctypes.data_as(C.POINTER(C.c_int32)),
                                  lu.ctypes.data_as(C.POINTER(C.c_double)))

    def _size(self, name):
        n3 = self.n ** 3
        return n3 * self.nfreq if name in PER_BIN else n3

    def upload_field(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32).reshape(-1)
        assert a.size == self._size(name)
        assert self._lib.orc_upload_field(self._h, FIELD_ID[name], a.ctypes.data_as(C.POINTER(C.c_float))) == 0

    def download_field(self, name):
        a = np.empty(self._size(name), dtype=np.float32)
        assert self._lib.orc_download_field(self._h, FIELD_ID[name], a.ctypes.data_as(C.POINTER(C.c_float))) == 0
        return a.reshape(-1, self.nfreq) if name in PER_BIN else a

    def set_time(self, t):
        self._lib.orc_set_time(self._h, float(t))

    def set_subgrid(self, sg):
        st, keep = make_subgrid(sg)
        assert self._lib.orc_set_subgrid(self._h, C.byref(st)) == 0

    def set_subgrid_delta_nl(self, d):
        a = np.ascontiguousarray(d, dtype=np.float32)
        assert self._lib.orc_set_subgrid_delta_nl(self._h, a.ctypes.data_as(C.POINTER(C.c_float))) == 0

    def download_subgrid_state(self):
        st, arrays = make_subgrid_state(self.n ** 3)
        assert self._lib.orc_download_subgrid_state(self._h, C.byref(st)) == 0
        return arrays

    def upload_subgrid_state(self, arrays):
        st, keep = make_subgrid_state(self.n ** 3, arrays)
        assert self._lib.orc_upload_subgrid_state(self._h, C.byref(st)) == 0

    def get_time(self):
        return self._lib.orc_get_time(self._h)

    def set_step(self, step):
        self._lib.orc_set_step(self._h, int(step))

    def set_opacity_dither(self, seed):
        self._lib.orc_set_opacity_dither(self._h, int(seed))

    def set_nphot_merge(self, v):
        self._lib.orc_set_nphot_merge(self._h, float(v))

    def init_rt(self):
        self._lib.orc_init_rt(self._h)

    def get_nphot_min_max(self):
        a, b = C.c_double(), C.c_double()
        self._lib.orc_get_nphot_min_max(self._h, C.byref(a), C.byref(b))
        return a.value, b.value

    def new_rays(self):
        self._lib.orc_new_rays(self._h)

    def update_rays(self):
        rc = self._lib.orc_update_rays(self._h)
        if rc:
            raise RuntimeError("oracle: path overflow")

    def find_avg_unu(self):
        self._lib.orc_find_avg_unu(self._h)

    def merge_rays(self):
        self._lib.orc_merge_rays(self._h)

    def split_rays(self):
        self._lib.orc_split_rays(self._h)

    def update_thermal(self):
        self._lib.orc_update_thermal(self._h)

    def step(self):
        st = FrtStepStats()
        rc = self._lib.orc_step(self._h, C.byref(st))
        if rc:
            raise RuntimeError("oracle: path overflow")
        return st.as_dict()

    def stats(self):
        st = FrtStepStats()
        self._lib.orc_get_stats(self._h, C.byref(st))
        return st.as_dict()

    def ray_count(self):
        return int(self._lib.orc_ray_count(self._h))

    def rays(self):
        n = self.ray_count()
        v, arrays = make_ray_view(n, self.nfreq)
        assert self._lib.orc_download_rays(self._h, C.byref(v)) == 0
        return {k: a[: v.count] for k, a in arrays.items()}

    def upload_rays(self, rays):
        n = len(rays["cell"])
        v, arrays = make_ray_view(n, self.nfreq, rays)
        self._keep = arrays
        assert self._lib.orc_upload_rays(self._h, C.byref(v)) == 0

    def update_cosmo_expansion(self, a_current, a_prev):
        self._lib.orc_update_cosmo_expansion(self._h, float(a_current), float(a_prev))

    def update_hydro_step(self, zz, a_rt, avg_rho, dt_cosmic):
        self._lib.orc_update_hydro_step(self._h, float(zz), float(a_rt), float(avg_rho), float(dt_cosmic))

    def otf_averages(self, avg_rho, temp_0=100.0):
        o = FrtOtf()
        self._lib.orc_otf_averages(self._h, float(avg_rho), float(temp_0), C.byref(o))
        return o.as_dict()

    def update_gadget_var(self):
        self._lib.orc_update_gadget_var(self._h)

    def gadget_average(self):
        self._lib.orc_gadget_average(self._h)

    def reset_gadget_var(self):
        self._lib.orc_reset_gadget_var(self._h)

    def gadget_step_time(self):
        return float(self._lib.orc_get_gadget_step_time(self._h))

    def set_dt(self, dt):
        self._lib.orc_set_dt(self._h, float(dt))

    def cell_size(self):
        return float(self._lib.orc_get_cell_size(self._h))

    def write_ray_file(self, path):
        rc = self._lib.orc_write_ray_file(self._h, str(path).encode())
        if rc:
            raise RuntimeError("oracle: cannot write %s" % path)

    def paths(self):
        n = self.ray_count()
        P = self.path_length
        end = np.zeros(n, dtype=np.int32)
        cells = np.zeros((n, P), dtype=np.int32)
        dists = np.zeros((n, P), dtype=np.float32)
        assert self._lib.orc_download_paths(self._h, n, end.ctypes.data_as(C.POINTER(C.c_int32)),
                                            cells.ctypes.data_as(C.POINTER(C.c_int32)),
                                            dists.ctypes.data_as(C.POINTER(C.c_float))) == 0
        return end, cells, dists

    def derived(self):
        out = (C.c_double * 17)()
        self._lib.orc_derived(self._h, out)
        return dict(dx=out[0], dx0=out[1], dt=out[2], clight=out[3], rmax=list(out[4:17]))


# ---- unit entry points -------------------------------------------------------------------
def dm(name, *args):
    return getattr(load(), "orc_dm_" + name)(*args)


def rate(name, x):
    return load().orc_rate(RATES[name], float(x))


def pix2vec(order, pix):
    out = (C.c_double * 3)()
    load().orc_pix2vec(int(order), int(pix), out)
    return np.array(out[:])


def vec2pix(order, v):
    return int(load().orc_vec2pix(int(order), float(v[0]), float(v[1]), float(v[2])))


def rotate(itheta, direction, v):
    a = (C.c_float * 3)(*[float(x) for x in v])
    o = (C.c_float * 3)()
    load().orc_rotate(int(itheta), int(direction), a, o)
    return np.array(o[:], dtype=np.float32)


def c2ray_cell(nH, nHe, avg_nh1, avg_nhe1, f_H1, f_He1, clump, rH2, rHe2, gamma_h, gamma_he, dt):
    out = (C.c_float * 4)()
    load().orc_c2ray_cell(nH, nHe, avg_nh1, avg_nhe1, f_H1, f_He1, clump, rH2, rHe2, gamma_h, gamma_he, dt, out)
    return np.array(out[:], dtype=np.float32)


def hubble(a, hh, om, ol):
    return load().orc_hubble(a, hh, om, ol)


def itheta_hash(seed, stream, counter):
    return int(load().orc_itheta_hash(seed, stream, counter))


def sg_deltasig(delta_nl, z):
    """deltaNL_to_deltasig of the checker (subgrid.cc:193-204)."""
    return float(load().orc_sg_deltasig(float(delta_nl), float(z)))
