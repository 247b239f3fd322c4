"""ctypes binding of the treecorr_amd C ABI (include/treecorr_amd.h).

The library is built in-tree (treecorr_amd/libtreecorr_amd.so, see build.py).
There is no fallback: if the library is missing or no GPU is usable, every
compute call raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, 'libtreecorr_amd.so')

TC_OK, TC_ERR_INVALID, TC_ERR_HIP, TC_ERR_UNSUPPORTED = 0, 1, 2, 3

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_ll = ctypes.c_longlong
_i = ctypes.c_int
_ip = ctypes.POINTER(ctypes.c_int)
_d = ctypes.c_double

# name -> (restype, argtypes)
SIGNATURES = {
    'tc_version': (_i, []),
    'tc_last_error': (ctypes.c_char_p, []),
    'tc_device_count': (_i, [ctypes.POINTER(_i)]),
    'tc_field_create': (_i, [_i, _i] + [_dp] * 7 + [_ll, _d, _d, _i, _ll, _i, _i, _i,
                                                     ctypes.POINTER(_vp)]),
    'tc_field_create_device': (_i, [_i, _i] + [_vp] * 7 + [_ll, _d, _d, _i, _ll, _i, _i, _i,
                                                            ctypes.POINTER(_vp)]),
    'tc_field_create_multi': (_i, [_i, _i] + [_dp] * 7 + [_ll, ctypes.POINTER(_ll), _i, _d, _d,
                                                          _i, _ll, _i, _i, _i, ctypes.POINTER(_vp)]),
    'tc_field_nobj': (_i, [_vp, ctypes.POINTER(_ll)]),
    'tc_field_size': (_i, [_vp, _dp]),
    'tc_field_ntop': (_i, [_vp, ctypes.POINTER(_ll)]),
    'tc_field_center': (_i, [_vp, _dp]),
    'tc_field_nnodes': (_i, [_vp, ctypes.POINTER(_ll)]),
    'tc_field_copy_nodes': (_i, [_vp, _vp, _vp]),
    'tc_field_destroy': (_i, [_vp]),
    'tc_corr2_create': (_i, [_i, _i, _i, _d, _d, _i, _d, _d, _d, _d, _d, _d, _d, _d] + [_dp] * 8 +
                        [ctypes.POINTER(_vp)]),
    'tc_corr2_process_auto': (_i, [_vp, _vp, _i, _i, _i]),
    'tc_corr2_process_cross': (_i, [_vp, _vp, _vp, _i, _i, _i]),
    'tc_corr2_process_pairs': (_i, [_vp, ctypes.POINTER(_vp), _i, ctypes.POINTER(_vp), _i, _ip,
                                    _ip, _i, _i, _i, _dp]),
    'tc_quick_assign': (_i, [_dp, _i, _dp, _dp, _dp, _ll, _ip]),
    'tc_select_patch': (_i, [_i, _dp, _i, _dp, _dp, _dp, _ll, _ip]),
    'tc_set_tree_speculation': (_i, [_i]),
    'tc_set_traversal_mode': (_i, [_i]),
    'tc_traversal_fallbacks': (_ll, []),
    'tc_persist_counters': (_i, [ctypes.POINTER(_ll)]),
    'tc_set_persist_params': (_i, [_i, _ll]),
    'tc_set_persist_grab': (_i, [_i]),
    'tc_set_graph_mode': (_i, [_i]),
    'tc_graph_stats': (_i, [ctypes.POINTER(_ll), ctypes.POINTER(_ll)]),
    'tc_cost_model_create': (_i, [ctypes.c_ulonglong, ctypes.POINTER(_vp)]),
    'tc_cost_model_destroy': (_i, [_vp]),
    'tc_cost_model_sizes': (_i, [ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    'tc_cost_model_params': (_i, [_vp, ctypes.POINTER(ctypes.c_float), _i]),
    'tc_cost_model_predict': (_i, [_vp, ctypes.POINTER(ctypes.c_float), _ll,
                                   ctypes.POINTER(ctypes.c_float)]),
    'tc_cost_model_grad': (_i, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                _ll, ctypes.POINTER(ctypes.c_float), _dp]),
    'tc_cost_model_train': (_i, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                 _ll, _i, _i, ctypes.c_float, _dp]),
    'tc_set_cost_model': (_i, [_vp]),
    'tc_cov_rowsums': (_i, [_i, _i, _i, _dp, _dp, _dp]),
    'tc_cov_gram': (_i, [_i, _i, _dp, _dp, _dp]),
    'tc_set_job_log': (_i, [_ll]),
    'tc_job_log_fetch': (_i, [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), _ll,
                              ctypes.POINTER(_ll)]),
    'tc_set_multipole_plain': (_i, [_i]),
    'tc_set_zeta_mode': (_i, [_i]),
    'tc_kmeans_run': (_i, [_dp, _i, _i, _dp, _dp, _dp, _dp, _ll, _i, _d, _d, _i,
                           ctypes.POINTER(ctypes.c_int)]),
    'tc_kmeans_run_alt': (_i, [_vp, _dp, _i, _dp, _dp, _dp, _dp, _ll, _i, _d, _d,
                               ctypes.POINTER(ctypes.c_int)]),
    'tc_kmeans_run_noalloc': (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _ll, _i, _d, _d, _i,
                                   ctypes.POINTER(ctypes.c_int)]),
    'tc_patch_split': (_i, [_ip, _ll, _i, _i, ctypes.POINTER(_dp), ctypes.POINTER(_dp),
                            ctypes.POINTER(_ll)]),
    'tc_corr2_trivially_zero': (_i, [_vp, _i, _i] + [_d] * 8 + [ctypes.POINTER(_i)]),
    'tc_corr2_counters': (_i, [_vp, ctypes.POINTER(_ll)]),
    'tc_corr2_destroy': (_i, [_vp]),
    'tc_generate_xyz': (_i, [_dp] * 6 + [_ll]),
    'tc_set_num_threads': (_i, [_i]),
    'tc_get_num_threads': (_i, [ctypes.POINTER(_i)]),
    'tc_set_timing': (_i, [_i]),
    'tc_copy_calibration': (_i, [_vp, _vp, _ll]),
    'tc_kernel_stats': (_i, [_i, _dp, ctypes.POINTER(_ll)]),
    'tc_reset_kernel_stats': (_i, []),
    'tc_set_traversal_chunk': (_i, [_ll]),
    'tc_corr3_create': (_i, [_i, _i, _d, _d, _i, _d, _d, _d, _i] + [_dp] * 19 +
                        [ctypes.POINTER(_vp)]),
    'tc_corr3_process_auto': (_i, [_vp, _vp, _i, _i, _i]),
    'tc_corr3_process_cross12': (_i, [_vp, _vp, _vp, _i, _i, _i, _i]),
    'tc_corr3_process_cross21': (_i, [_vp, _vp, _vp, _i, _i, _i, _i]),
    'tc_corr3_process_cross': (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i]),
    'tc_corr3_counters': (_i, [_vp, ctypes.POINTER(_ll)]),
    'tc_corr3_zeta_cells': (_i, [_vp, ctypes.POINTER(_ll)]),
    'tc_corr3_trivially_zero': (_i, [_vp, _i] + [_d] * 12 + [_i, _i, ctypes.POINTER(_i)]),
    'tc_corr3_destroy': (_i, [_vp]),
}

_LIB = None


class TreeCorrAMDError(RuntimeError):
    """Raised when a call into the GPU engine fails."""


def lib():
    """Load the shared library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise TreeCorrAMDError(
                f'{LIBPATH} is missing: build it with `python -m treecorr_amd.build`')
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def exported_symbols():
    """Names from SIGNATURES that the loaded library exports."""
    L = lib()
    return [n for n in SIGNATURES if hasattr(L, n)]


def check(rc):
    if rc != TC_OK:
        msg = lib().tc_last_error().decode(errors='replace')
        if rc == TC_ERR_UNSUPPORTED:
            raise NotImplementedError(msg)
        if rc == TC_ERR_INVALID:
            raise ValueError(msg)
        raise TreeCorrAMDError(msg)


def dptr(a):
    """double* for a contiguous float64 array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def device_count():
    n = ctypes.c_int(0)
    lib().tc_device_count(ctypes.byref(n))
    return n.value
