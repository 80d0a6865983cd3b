"""Loader of the HIP library ``baylorlqcd_amd/libqqcdhip.so``.

The library is built in-tree (``make -C baylorlqcd_amd/csrc``) and loaded
from this folder only.  There is no fallback: if the library or a GPU is
missing, every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqqcdhip.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_dpp = ctypes.POINTER(_dp)
_L = None

# (name, argtypes) of every exported entry point of include/qqcd_hip.h
SIGNATURES = {
    "qqcd_init": [_ip, _ip, ctypes.c_int],
    "qqcd_finalize": [],
    "qqcd_last_error": [],
    "qqcd_ref_sizes": [_lp, _lp],
    "qqcd_gauge_load": [_dp, _dp],
    "qqcd_gauge_load_canonical": [_dp, _dp],
    "qqcd_hsingle": [_dp, _dp, _dp, ctypes.c_int, _dp, _ip, ctypes.c_int],
    "qqcd_hdouble": [_dp, _dp, _dp, ctypes.c_int, _dp, ctypes.c_double, _ip],
    "qqcd_hdbletm": [_dp, _dp, _dp, ctypes.c_int, _dp, _dp, ctypes.c_int, _ip],
    "qqcd_gammamult": [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "qqcd_gam5x": [_dp],
    "qqcd_vecdot": [_dp, _dp, _dp],
    "qqcd_pointsource": [_dp, _dp, _ip, ctypes.c_int, ctypes.c_int],
    "qqcd_gmresdr": [_dp, _dp, _ip, ctypes.c_double, ctypes.c_int, _ip, _dp, _dp, _dp, ctypes.c_int, _ip],
    "qqcd_gmresproject": [_dp, _dp, ctypes.c_int, _ip, ctypes.c_double, ctypes.c_int, _ip, _dp, _dp, _dp,
                          ctypes.c_int, _ip, ctypes.c_int, _ip, _dp, ctypes.c_int],
    "qqcd_psc": [_dp, _dp, _ip, ctypes.c_double, ctypes.c_int, _ip, _dp, _dp, _dp, ctypes.c_int, _ip,
                 ctypes.c_int, ctypes.c_double],
    "qqcd_bicgstab": [_dp, _dp, ctypes.c_double, ctypes.c_int, _ip, _dp, _dp, _dp, ctypes.c_int, _ip],
    "qqcd_fermprop": [_dp, _dp, _dp, ctypes.c_int, _dp, _ip, ctypes.c_double, ctypes.c_int, _ip, ctypes.c_int,
                      _dp, _dp, _dp, ctypes.c_int, _ip],
    "qqcd_bench_dslash": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp],
    "qqcd_bench_dslash_rot": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp],
    "qqcd_hop_canonical": [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "qqcd_schur_canonical": [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "qqcd_set_xcd_remap": [ctypes.c_int],
    "qqcd_set_dslash_tune": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "qqcd_set_multi_tune": [ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "qqcd_set_multi_lds": [ctypes.c_int],
    "qqcd_set_arnoldi_steps": [ctypes.c_int],
    "qqcd_set_pair_norms": [ctypes.c_int],
    "qqcd_set_batch_mid": [ctypes.c_int],
    "qqcd_set_step_graphs": [ctypes.c_int],
    "qqcd_gauge_random": [ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong, _dp],
    "qqcd_gauge_get_canonical": [_dp],
    "qqcd_gauge_heatbath": [ctypes.c_double, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_int, _dp],
    "qqcd_gauge_plaquette": [_dp],
    "qqcd_last_solve_stats": [_dp, _ip, _ip, _ip, _ip],
    "qqcd_deflation_clear": [],
    "qqcd_solve_mixed": [_dp, _dp, _ip, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                         ctypes.c_int, _dp, _dp, _dp, _ip, _ip],
    "qqcd_bench_solve": [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, _dp, _ip, ctypes.c_double, ctypes.c_int,
                         ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _ip, _ip, _dp],
    "qqcd_bench_copy": [ctypes.c_double, ctypes.c_int, _dp],
    "qqcd_bench_blas": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp],
    "qqcd_dense_schur": [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp],
    "qqcd_dense_reorder": [ctypes.c_int, _dp, _dp, ctypes.c_int],
    "qqcd_dense_lsq": [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp],
    "qqcd_dense_harmonic_ritz": [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp],
    "qqcd_z4source": [_dp, _dp, ctypes.c_ulonglong, ctypes.c_long, ctypes.c_double],
    "qqcd_z2source": [_dp, _dp, ctypes.c_ulonglong, ctypes.c_long, ctypes.c_double],
    "qqcd_fermprop_multi": [ctypes.c_int, _dpp, _dpp, _dp, _dp, _ip, ctypes.c_double, _ip, _dp, _dpp, _dpp, _ip,
                            _dp],
    "qqcd_noise_loop": [ctypes.c_int, ctypes.c_ulonglong, ctypes.c_double, _dp, _dp, _ip, _dp, _ip,
                        ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, _dp, _ip, _dp,
                        _ip],
    "qqcd_noise_loop_lagfib": [ctypes.c_int, _ip, _ip, _ip, ctypes.c_double, _dp, _dp, _ip, _dp, _ip,
                               ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, _dp, _ip,
                               _dp, _ip],
    "qqcd_loopops": [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _dp, _dp],
    "qqcd_lagfib_init": [ctypes.c_int, ctypes.c_int, _ip, _ip, _ip],
    "qqcd_rnds": [ctypes.c_long, _dp, _ip, _ip, _ip],
    "qqcd_rnds_noalloc": [ctypes.c_long, _dp, _ip, _ip, _ip],
    "qqcd_z4source_lagfib": [_dp, _dp, _ip, _ip, _ip],
    "qqcd_z2source_lagfib": [_dp, _dp, _ip, _ip, _ip],
    "qqcd_sweep": [_ip, _dp, ctypes.c_int, ctypes.c_int, _dp, _ip, _ip, _ip],
    "qqcd_smear": [_dp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, _ip],
    "qqcd_bench_smear": [ctypes.c_int, ctypes.c_int, _dp],
    "qqcd_gauge_get_ref": [_dp],
    "qqcd_solve_log": [ctypes.c_int, _ip, ctypes.c_char_p, _ip, _dp],
    "qqcd_bench_lagfib": [ctypes.c_long, ctypes.c_int, ctypes.c_long, _dp],
    "qqcd_twmshift": [_dp, _dp, _dp, ctypes.c_int, _dp, _dp, ctypes.c_int, _ip],
    "qqcd_gamma5mult": [_dp, _dp, _dp, ctypes.c_int, _dp, _dp, ctypes.c_int, _ip],
    "qqcd_gfwrite": [_ip, _dp, ctypes.c_char_p, _ip, ctypes.c_double, _dp, _dp, ctypes.c_int, ctypes.c_int],
    "qqcd_gfread": [_ip, _dp, ctypes.c_char_p, _ip, _dp, _dp, _dp, _ip, _ip],
    "qqcd_fpwrite": [_ip, _dp, _dp, ctypes.c_char_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                     ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "qqcd_fpread": [_ip, _dp, _dp, ctypes.c_char_p, _dp, _dp, _dp, _ip, _ip, _ip, _ip],
    "qqcd_propagator": [ctypes.c_int, _ip, _dp, _ip, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                        ctypes.c_void_p, ctypes.c_void_p, _ip, _dp, _dp, _ip],
    "qqcd_loopsum": [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _ip, ctypes.c_double, _dp],
    "qqcd_trace16": [_dp, _dp, _dp, _dp, _dp],
    "qqcd_batch_set": [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                       ctypes.c_int, ctypes.c_int],
    "qqcd_batch_stats": [_ip, _ip, _lp, _lp, _ip, _ip],
    "qqcd_batch_rejected": [_ip],
    "qqcd_mixed_stats": [_ip, _ip, _ip, _ip, _ip, _dp],
    "qqcd_set_fault": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "qqcd_hop_multi_canonical": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp],
    "qqcd_schur_multi_canonical": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                   ctypes.c_int, ctypes.c_int, _dp, _dp],
    "qqcd_bench_dslash_multi": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp],
    "qqcd_autotune_multi": [ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip, _dp],
    "qqcd_nn_init": [ctypes.c_int, _ip, ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong],
    "qqcd_nn_set_weights": [ctypes.c_int, _ip, _dp],
    "qqcd_nn_get_weights": [_ip, _ip, _dp],
    "qqcd_nn_layer_canonical": [ctypes.c_int, ctypes.c_int, _dp, _dp],
    "qqcd_nn_wgrad_canonical": [ctypes.c_int, _dp, _dp, _dp],
    "qqcd_nn_apply_canonical": [_dp, _dp],
    "qqcd_nn_train": [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                      ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int,
                      ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong, _dp, _dp],
    "qqcd_bench_nn": [ctypes.c_int, ctypes.c_int, _dp],
    "qqcd_bench_learned": [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, _dp, _ip, ctypes.c_double, ctypes.c_int,
                           ctypes.c_int, _dp, _ip, _ip, _dp, _ip, _ip],
}


class QQCDError(RuntimeError):
    pass


def lib():
    """The loaded library; raises if it was not built."""
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise QQCDError(f"{LIB_PATH} is missing: build it with `make -C baylorlqcd_amd/csrc` "
                            "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.argtypes = args
            f.restype = ctypes.c_char_p if name == "qqcd_last_error" else ctypes.c_int
        _L = L
    return _L


def check(rc):
    if rc != 0:
        raise QQCDError(lib().qqcd_last_error().decode())
    return rc


def dp(a):
    if a is None:
        return ctypes.cast(None, _dp)
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and
            (a.flags.c_contiguous or a.flags.f_contiguous)):
        raise TypeError("expected a contiguous float64 array")
    return a.ctypes.data_as(_dp)


def ip(a):
    if a is None:
        return ctypes.cast(None, _ip)
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a.ctypes.data_as(_ip), a


def dpp(arrays):
    """A C array of double* for a list of contiguous float64 arrays."""
    ptrs = (_dp * len(arrays))(*[dp(a) for a in arrays])
    return ctypes.cast(ptrs, _dpp), ptrs


def ints(v):
    return np.ascontiguousarray(v, dtype=np.int32)
