"""Loader and prototypes of libflexrt_gpu.so (the C ABI in include/flexrt_gpu.h).

The library is loaded from the package directory only.  If it is absent the
import of any GPU entry point raises: there is no CPU fallback.
"""
import ctypes as C
import os

from ._abi import FrtConfig, FrtFileExtras, FrtOtf, FrtRayView, FrtStepStats, FrtSubgrid, FrtSubgridState
from ._build import LIB_PATH

_lib = None

_CTX = C.c_void_p
_PROTOS = {
    "frt_version": (C.c_char_p, []),
    "frt_config_default": (C.c_int, [C.POINTER(FrtConfig), C.c_int32]),
    "frt_create": (C.c_int, [C.POINTER(FrtConfig), C.POINTER(_CTX)]),
    "frt_destroy": (C.c_int, [_CTX]),
    "frt_last_error": (C.c_char_p, [_CTX]),
    "frt_set_spectrum": (C.c_int, [_CTX, C.POINTER(C.c_double), C.POINTER(C.c_float)]),
    "frt_set_sources": (C.c_int, [_CTX, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "frt_upload_field": (C.c_int, [_CTX, C.c_int32, C.POINTER(C.c_float)]),
    "frt_download_field": (C.c_int, [_CTX, C.c_int32, C.POINTER(C.c_float)]),
    "frt_set_time": (C.c_int, [_CTX, C.c_float]),
    "frt_set_subgrid": (C.c_int, [_CTX, C.POINTER(FrtSubgrid)]),
    "frt_set_subgrid_delta_nl": (C.c_int, [_CTX, C.POINTER(C.c_float)]),
    "frt_download_subgrid_state": (C.c_int, [_CTX, C.POINTER(FrtSubgridState)]),
    "frt_upload_subgrid_state": (C.c_int, [_CTX, C.POINTER(FrtSubgridState)]),
    "frt_get_time": (C.c_int, [_CTX, C.POINTER(C.c_float)]),
    "frt_set_step": (C.c_int, [_CTX, C.c_int64]),
    "frt_set_dt": (C.c_int, [_CTX, C.c_float]),
    "frt_get_cell_size": (C.c_int, [_CTX, C.POINTER(C.c_float)]),
    "frt_update_cosmo_expansion": (C.c_int, [_CTX, C.c_float, C.c_float]),
    "frt_set_opacity_dither": (C.c_int, [_CTX, C.c_uint64]),
    "frt_update_hydro_step": (C.c_int, [_CTX, C.c_float, C.c_float, C.c_float, C.c_float]),
    "frt_otf_averages": (C.c_int, [_CTX, C.c_float, C.c_float, C.POINTER(FrtOtf)]),
    "frt_otf_averages_noalloc": (C.c_int, [_CTX, C.c_float, C.c_float, C.POINTER(FrtOtf)]),
    "frt_write_otf": (C.c_int, [_CTX, C.c_char_p, C.c_float, C.c_float]),
    "frt_init_rt": (C.c_int, [_CTX]),
    "frt_get_nphot_min_max": (C.c_int, [_CTX, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "frt_new_rays": (C.c_int, [_CTX]),
    "frt_update_rays": (C.c_int, [_CTX]),
    "frt_find_avg_unu": (C.c_int, [_CTX]),
    "frt_merge_rays": (C.c_int, [_CTX]),
    "frt_split_rays": (C.c_int, [_CTX]),
    "frt_update_thermal": (C.c_int, [_CTX]),
    "frt_step": (C.c_int, [_CTX, C.POINTER(FrtStepStats)]),
    "frt_get_stats": (C.c_int, [_CTX, C.POINTER(FrtStepStats)]),
    "frt_set_step_graph": (C.c_int, [_CTX, C.c_int32]),
    "frt_set_tuning": (C.c_int, [_CTX, C.c_char_p, C.c_int64]),
    "frt_get_tuning": (C.c_int, [_CTX, C.c_char_p, C.POINTER(C.c_int64)]),
    "frt_device_arch": (C.c_int, [C.c_char_p, C.c_int32]),
    "frt_selftest_scan_sort": (C.c_int, [_CTX, C.c_int64, C.c_int32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                         C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int64)]),
    "frt_selftest_sort_cells": (C.c_int, [_CTX, C.c_int64, C.POINTER(C.c_uint32), C.POINTER(C.c_int32),
                                          C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "frt_ray_count": (C.c_int, [_CTX, C.POINTER(C.c_int64)]),
    "frt_download_rays": (C.c_int, [_CTX, C.POINTER(FrtRayView)]),
    "frt_upload_rays": (C.c_int, [_CTX, C.POINTER(FrtRayView)]),
    "frt_download_paths": (C.c_int, [_CTX, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_float)]),
    "frt_write_vif_file": (C.c_int, [_CTX, C.c_char_p]),
    "frt_update_gadget_var": (C.c_int, [_CTX]),
    "frt_write_gadget_file": (C.c_int, [_CTX, C.c_char_p]),
    "frt_reset_gadget_var": (C.c_int, [_CTX]),
    "frt_get_gadget_step_time": (C.c_int, [_CTX, C.POINTER(C.c_float)]),
    "frt_adj_enable": (C.c_int, [_CTX, C.c_int32]),
    "frt_adj_ray_count": (C.c_int, [_CTX, C.POINTER(C.c_int64)]),
    "frt_adj_find_avg_unu": (C.c_int, [_CTX] + [C.POINTER(C.c_double)] * 7),
    "frt_checkpoint_save": (C.c_int, [_CTX, C.POINTER(C.c_void_p)]),
    "frt_checkpoint_restore": (C.c_int, [_CTX, C.c_void_p]),
    "frt_checkpoint_free": (C.c_int, [C.c_void_p]),
    "frt_write_ray_file": (C.c_int, [_CTX, C.c_char_p]),
    "frt_read_ray_file": (C.c_int, [_CTX, C.c_char_p]),
    "frt_write_gas_file": (C.c_int, [_CTX, C.c_char_p, C.POINTER(FrtFileExtras)]),
    "frt_read_gas_file": (C.c_int, [_CTX, C.c_char_p, C.POINTER(FrtFileExtras)]),
    "frt_write_treion_file": (C.c_int, [_CTX, C.c_char_p, C.POINTER(FrtFileExtras)]),
    "frt_read_treion_file": (C.c_int, [_CTX, C.c_char_p, C.POINTER(FrtFileExtras)]),
    "frt_adj_run_begin": (C.c_int, [_CTX]),
    "frt_adj_seed": (C.c_int, [_CTX, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "frt_adj_step_reverse": (C.c_int, [_CTX]),
    "frt_adj_run_finish": (C.c_int, [_CTX, C.c_float, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "frt_set_profiling": (C.c_int, [_CTX, C.c_int32]),
    "frt_kernel_time": (C.c_int, [_CTX, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "frt_reset_kernel_times": (C.c_int, [_CTX]),
    "frt_deposit_bytes": (C.c_int, [_CTX, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "frt_copy_kernel": (C.c_int, [_CTX, C.c_int64, C.c_int32]),
    "frt_synchronize": (C.c_int, [_CTX]),
}

EXPORTED = sorted(_PROTOS)


def lib_path():
    return LIB_PATH


def load():
    """Load the built library (raises FileNotFoundError if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            "libflexrt_gpu.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
