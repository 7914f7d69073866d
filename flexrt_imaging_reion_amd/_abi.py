"""ctypes mirror of the structs in include/flexrt_gpu.h.

Shared by the GPU binding (``_lib``) and the CPU checker binding, which both
take the same ``frt_config`` / ``frt_step_stats`` / ``frt_ray_view`` structs.
"""
import ctypes as C

import numpy as np

FRT_MAX_FREQ = 8
FRT_NTHETA = 42875

FRT_OK = 0
FRT_E_ARG = 1
FRT_E_HIP = 2
FRT_E_CAPACITY = 3
FRT_E_PATH_OVERFLOW = 4
FRT_E_STATE = 5
FRT_E_NOMEM = 6
FRT_E_IO = 7

PREC_FP64 = 0
PREC_MIXED = 1

# Field ids, in the order of frt_field (global_variables.cc names).
FIELDS = [
    "nH", "nHe", "nH1", "nHe1", "nH2", "nHe2", "ne", "n_tot", "f_H1", "f_He1", "f_H2", "f_He2",
    "temp", "dne_dt", "clump", "nH1_prev", "avg_nh1", "avg_nhe1", "recomb_H2", "recomb_He2",
    "gamma_H1_tot", "gamma_He1_tot", "recrate_H2", "recrate_He2", "treion", "treion_flag",
    "heat_rate", "cool_rate", "IF_speed", "Inc_Flux", "gamma_H1_nu", "gamma_He1_nu", "ngam_nu",
    "rho", "f_He3", "mfp", "mfp_912", "avg_gamma_prev", "gadget_gamma", "gadget_heat", "ddelta_dt", "rho_prev",
    "nHe3", "nH2_prev", "nHe1_prev", "nHe2_prev", "nHe3_prev", "recomb_He3", "u_nu",
]
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}
PER_BIN = {"gamma_H1_nu", "gamma_He1_nu", "ngam_nu", "u_nu"}


class FrtConfig(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("nfreq", C.c_int32), ("path_length", C.c_int32), ("itercount", C.c_int32),
        ("precision", C.c_int32), ("hpx_lvl", C.c_int32), ("l_max", C.c_int32),
        ("ray_per_cell_max", C.c_int32), ("rec_case_b", C.c_int32), ("temp_ev", C.c_int32),
        ("recomb_cool", C.c_int32), ("free_free", C.c_int32), ("compton", C.c_int32),
        ("frac_limit", C.c_int32), ("subgrid", C.c_int32), ("c2ray_off", C.c_int32),
        ("ray_capacity", C.c_int64), ("seed", C.c_uint64),
        ("box_ckpc", C.c_double), ("a_rt", C.c_double), ("zz", C.c_double), ("dt", C.c_double),
        ("cl_factor", C.c_double), ("nl_const", C.c_double), ("safety_factor", C.c_double),
        ("min_dist", C.c_double), ("temp_min", C.c_double), ("hh", C.c_double),
        ("omega_m", C.c_double), ("omega_l", C.c_double),
    ]


class FrtStepStats(C.Structure):
    _fields_ = [
        ("step", C.c_int64), ("rays_start", C.c_int64), ("rays_moved", C.c_int64),
        ("segments", C.c_int64), ("rays_dumped", C.c_int64), ("rays_terminated", C.c_int64),
        ("rays_counted", C.c_int64), ("num_thresh", C.c_int64), ("merge_candidates", C.c_int64),
        ("merge_groups", C.c_int64), ("rays_merged_away", C.c_int64), ("rays_split", C.c_int64),
        ("rays_new", C.c_int64), ("rays_end", C.c_int64), ("thermal_substeps", C.c_int64),
        ("nphot_min", C.c_double), ("nphot_max", C.c_double), ("nphot_merge", C.c_float),
        ("t", C.c_float),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class FrtRayView(C.Structure):
    _fields_ = [
        ("count", C.c_int64), ("order", C.POINTER(C.c_int16)), ("pixel", C.POINTER(C.c_int32)),
        ("itheta", C.POINTER(C.c_uint16)), ("xhat", C.POINTER(C.c_float)),
        ("yhat", C.POINTER(C.c_float)), ("zhat", C.POINTER(C.c_float)),
        ("nphot", C.POINTER(C.c_float)), ("r", C.POINTER(C.c_float)), ("dr", C.POINTER(C.c_float)),
        ("x", C.POINTER(C.c_float)), ("y", C.POINTER(C.c_float)), ("z", C.POINTER(C.c_float)),
        ("time", C.POINTER(C.c_float)), ("cell", C.POINTER(C.c_int32)),
    ]


# frt_file_extras: per-cell fields of the restart files the library does not
# evolve, in struct order.
EXTRA_FLOAT_FIELDS = ["rho", "f_He3", "mfp", "mfp_912", "fion", "tcross", "treion_out"]
EXTRA_FLAG_FIELDS = ["tcross_flag", "neutral_flag"]


# write_otf's averages (frt_otf): (name, values); float fields except npht, nsrc.
OTF_FIELDS = [("fH1", 2), ("fH2", 2), ("fHe1", 2), ("fHe2", 2), ("fHe3", 2), ("nH1", 1), ("nHe1", 1),
              ("nH", 3), ("nHe", 3), ("ne", 1), ("GH1", 7), ("RH2", 5), ("T", 5), ("MFP", 6),
              ("MFP_912", 4), ("Heat", 5), ("Cool", 5)]


class FrtOtf(C.Structure):
    _fields_ = [(n, C.c_float * k if k > 1 else C.c_float) for n, k in OTF_FIELDS] + \
               [("npht", C.c_double), ("nsrc", C.c_int32), ("fion", C.c_float * 2)]

    def as_dict(self):
        d = {}
        for n, k in OTF_FIELDS:
            v = getattr(self, n)
            d[n] = list(v) if k > 1 else v
        d["npht"] = self.npht
        d["nsrc"] = self.nsrc
        d["fion"] = list(self.fion)
        return d


class FrtFileExtras(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_float)) for n in EXTRA_FLOAT_FIELDS] + \
               [(n, C.POINTER(C.c_uint8)) for n in EXTRA_FLAG_FIELDS]


def make_extras(n3, arrays=None, alloc=False):
    """An frt_file_extras over numpy arrays: the given ones (write), or fresh
    zero arrays for every field (alloc=True, read).  Returns (struct, arrays)."""
    arrays = dict(arrays or {})
    if alloc:
        for n in EXTRA_FLOAT_FIELDS:
            arrays.setdefault(n, np.zeros(n3, dtype=np.float32))
        for n in EXTRA_FLAG_FIELDS:
            arrays.setdefault(n, np.zeros(n3, dtype=np.uint8))
    ex = FrtFileExtras()
    for n in EXTRA_FLOAT_FIELDS:
        if n in arrays:
            arrays[n] = np.ascontiguousarray(arrays[n], dtype=np.float32).reshape(-1)
            assert arrays[n].size == n3, n
            setattr(ex, n, arrays[n].ctypes.data_as(C.POINTER(C.c_float)))
    for n in EXTRA_FLAG_FIELDS:
        if n in arrays:
            arrays[n] = np.ascontiguousarray(arrays[n], dtype=np.uint8).reshape(-1)
            assert arrays[n].size == n3, n
            setattr(ex, n, arrays[n].ctypes.data_as(C.POINTER(C.c_uint8)))
    return ex, arrays


RAY_FIELDS = [
    ("order", np.int16), ("pixel", np.int32), ("itheta", np.uint16), ("xhat", np.float32),
    ("yhat", np.float32), ("zhat", np.float32), ("nphot", np.float32), ("r", np.float32),
    ("dr", np.float32), ("x", np.float32), ("y", np.float32), ("z", np.float32),
    ("time", np.float32), ("cell", np.int32),
]


def make_ray_view(count, nfreq, arrays=None):
    """Allocate (or wrap) numpy arrays for ``count`` rays and a view on them."""
    if arrays is None:
        arrays = {}
        for name, dt in RAY_FIELDS:
            shape = (count, nfreq) if name == "nphot" else (count,)
            arrays[name] = np.zeros(shape, dtype=dt)
    else:
        arrays = {name: np.ascontiguousarray(arrays[name], dtype=dt) for name, dt in RAY_FIELDS}
    v = FrtRayView()
    v.count = count
    for name, dt in RAY_FIELDS:
        ctype = np.ctypeslib.as_ctypes_type(dt)
        setattr(v, name, arrays[name].ctypes.data_as(C.POINTER(ctype)))
    return v, arrays


def config_struct(cfg):
    """Build an FrtConfig from a dict of field values."""
    s = FrtConfig()
    for name, _ in FrtConfig._fields_:
        if name in cfg:
            setattr(s, name, cfg[name])
    return s


# ---- sub-grid I-front model (frt_subgrid, frt_subgrid_state) ----------------------------
NTCROSS = 10
SUBGRID_STATE = [("fion", 1), ("fion_max", 1), ("tcross", 1), ("tcross_flag", 1), ("treion_out", 1),
                 ("delta_fion", 1), ("tcross_new", NTCROSS), ("tcross_check", NTCROSS), ("fh1_nodes", 5)]


class FrtSubgrid(C.Structure):
    _fields_ = [("nt", C.c_int32), ("reserved", C.c_int32), ("time_rh", C.POINTER(C.c_float)),
                ("fh1_rh", C.POINTER(C.c_float)), ("tre_dom", C.c_float * 3), ("t_z12", C.c_float),
                ("t_z8", C.c_float), ("tinit", C.c_float), ("avg_rho", C.c_float), ("temp_0", C.c_float),
                ("fh1_0", C.c_float), ("delta_nl", C.c_float * 3)]


class FrtSubgridState(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_float)) for n, _ in SUBGRID_STATE]


def make_subgrid(sg):
    """An frt_subgrid over a dict (see subgrid.make_params); returns (struct, arrays kept alive)."""
    keep = {k: np.ascontiguousarray(sg[k], dtype=np.float32).reshape(-1) for k in ("time_rh", "fh1_rh")}
    nt = int(sg["nt"])
    assert keep["time_rh"].size == 27 * nt and keep["fh1_rh"].size == 27 * nt
    s = FrtSubgrid()
    s.nt = nt
    s.time_rh = keep["time_rh"].ctypes.data_as(C.POINTER(C.c_float))
    s.fh1_rh = keep["fh1_rh"].ctypes.data_as(C.POINTER(C.c_float))
    for i in range(3):
        s.tre_dom[i] = float(sg["tre_dom"][i])
        s.delta_nl[i] = float(sg["delta_nl"][i])
    for k in ("t_z12", "t_z8", "tinit", "avg_rho", "temp_0", "fh1_0"):
        setattr(s, k, float(sg[k]))
    return s, keep


def make_subgrid_state(n3, arrays=None):
    """An frt_subgrid_state over numpy arrays: the given ones (upload), or
    fresh arrays for every member (download).  Returns (struct, arrays)."""
    if arrays is None:
        arrays = {n: np.zeros(n3 * w, dtype=np.float32) for n, w in SUBGRID_STATE}
    else:
        arrays = {n: np.ascontiguousarray(arrays[n], dtype=np.float32).reshape(-1) for n, _ in SUBGRID_STATE
                  if n in arrays}
    st = FrtSubgridState()
    for n, w in SUBGRID_STATE:
        if n in arrays:
            assert arrays[n].size == n3 * w, n
            setattr(st, n, arrays[n].ctypes.data_as(C.POINTER(C.c_float)))
    return st, arrays
