"""ctypes binding of libmqct_hip.so (the C ABI declared in include/mqct_hip.h).

The library is built in-tree by mqct_amd/build.py.  Importing this module
never falls back to anything else: if the library is missing or cannot be
loaded, LibraryMissing is raised with the reason.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libmqct_hip.so")

MQCT_PREC_FP64, MQCT_PREC_FP32, MQCT_PREC_BF16X3, MQCT_PREC_SURROGATE = 0, 1, 2, 3
MQCT_PES_LEGENDRE, MQCT_PES_REALSH, MQCT_PES_SITES = 0, 1, 2
MQCT_PROP_CC_RK4, MQCT_PROP_CC_ODEINT, MQCT_PROP_AT_RK4, MQCT_PROP_AT_ODEINT = 0, 1, 2, 3


class LibraryMissing(RuntimeError):
    pass


class MqctError(RuntimeError):
    """A library call returned a non-zero status; the message is the
    library's own (the reference would STOP with it)."""


c_int_p = ctypes.POINTER(ctypes.c_int)
c_dbl_p = ctypes.POINTER(ctypes.c_double)
c_flt_p = ctypes.POINTER(ctypes.c_float)
c_u8_p = ctypes.POINTER(ctypes.c_uint8)
c_i64_p = ctypes.POINTER(ctypes.c_int64)


class GridT(ctypes.Structure):
    _fields_ = [
        ("n_beta", ctypes.c_int), ("xb", c_dbl_p), ("wb", c_dbl_p), ("xbm", ctypes.c_double), ("xbl", ctypes.c_double),
        ("n_gamma", ctypes.c_int), ("xg", c_dbl_p), ("wg", c_dbl_p), ("xgm", ctypes.c_double), ("xgl", ctypes.c_double),
        ("n_alpha2", ctypes.c_int), ("xaa", c_dbl_p), ("waa", c_dbl_p), ("xaam", ctypes.c_double), ("xaal", ctypes.c_double),
        ("n_beta2", ctypes.c_int), ("xbb", c_dbl_p), ("wbb", c_dbl_p), ("xbbm", ctypes.c_double), ("xbbl", ctypes.c_double),
        ("n_gamma2", ctypes.c_int), ("xgg", c_dbl_p), ("wgg", c_dbl_p), ("xggm", ctypes.c_double), ("xggl", ctypes.c_double),
        ("eq_gamma", ctypes.c_int),
        ("ref_xz", ctypes.c_int),
    ]


class TrajT(ctypes.Structure):
    _fields_ = [
        ("s_st", ctypes.c_int), ("at_path", ctypes.c_int),
        ("l_real", ctypes.c_double), ("E_coll", ctypes.c_double),
        ("dt", ctypes.c_double), ("eps", ctypes.c_double),
        ("hmin", ctypes.c_double), ("R_start", ctypes.c_double),
        ("time_lim", ctypes.c_double),
        ("numb_oscl_prds", ctypes.c_int), ("numb_orb_prds", ctypes.c_int),
        ("max_steps", ctypes.c_int),
    ]


class TrajStatT(ctypes.Structure):
    _fields_ = [
        ("t_end", ctypes.c_double), ("n_steps", ctypes.c_int),
        ("n_rhs", ctypes.c_int), ("fail", ctypes.c_int),
        ("norm_err", ctypes.c_double),
    ]


# name -> (restype, argtypes); the complete export list of include/mqct_hip.h
SIGNATURES = {
    "mqct_hip_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "mqct_hip_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "mqct_hip_last_error": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "mqct_hip_version": (ctypes.c_int, []),
    "mqct_hip_set_basis": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_int_p, c_int_p, c_int_p,
                                          ctypes.c_int, c_int_p, c_int_p, ctypes.c_int, c_dbl_p, c_dbl_p]),
    "mqct_hip_set_grid": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(GridT)]),
    "mqct_hip_set_vgrid": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_dbl_p, c_dbl_p]),
    "mqct_hip_eval_pes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_dbl_p, ctypes.c_int, ctypes.c_int, c_dbl_p]),
    "mqct_hip_matrix_build": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, c_int_p, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int, ctypes.c_int, c_dbl_p, c_dbl_p, c_u8_p]),
    "mqct_hip_matrix_get": (ctypes.c_int, [ctypes.c_void_p, c_dbl_p, c_dbl_p]),
    "mqct_hip_build_timing": (ctypes.c_int, [ctypes.c_void_p, c_dbl_p]),
    "mqct_hip_set_mij_shift": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "mqct_hip_mij_spline": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_dbl_p, ctypes.c_int64, c_dbl_p,
                                           c_dbl_p]),
    "mqct_hip_write_mij_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_int64, ctypes.c_int, c_int_p, c_int_p,
                                               c_int_p, c_int_p, c_int_p, c_int_p, c_dbl_p]),
    "mqct_hip_write_mij_dir": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                              ctypes.c_int, c_int_p, c_dbl_p, ctypes.c_double, ctypes.c_int,
                                              ctypes.c_int]),
    "mqct_hip_read_mij_dir": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                             c_i64_p, c_dbl_p, c_int_p, c_dbl_p]),
    "mqct_hip_set_progress": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "mqct_hip_set_identical": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_int_p, c_int_p, c_int_p]),
    "mqct_hip_set_expansion_terms": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_int_p]),
    "mqct_hip_expansion_project": (ctypes.c_int, [ctypes.c_void_p, c_dbl_p]),
    "mqct_hip_expansion_build": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_dbl_p, c_dbl_p, ctypes.c_int64,
                                                c_int_p, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                                ctypes.c_int, c_dbl_p, c_dbl_p, c_u8_p]),
    "mqct_hip_expansion_timing": (ctypes.c_int, [ctypes.c_void_p, c_dbl_p]),
    "mqct_hip_load_matrix": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_dbl_p, c_int_p, c_int_p, ctypes.c_int64,
                                            c_int_p, ctypes.c_int, c_dbl_p, c_dbl_p, c_dbl_p, ctypes.c_double,
                                            ctypes.c_int]),
    "mqct_hip_use_built_matrix": (ctypes.c_int, [ctypes.c_void_p, c_dbl_p, ctypes.c_double, ctypes.c_int]),
    "mqct_hip_load_paths": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_i64_p, c_dbl_p]),
    "mqct_hip_propagate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(TrajT),
                                          ctypes.c_int, c_dbl_p, ctypes.POINTER(TrajStatT)]),
    "mqct_hip_propagate_timing": (ctypes.c_int, [ctypes.c_void_p, c_dbl_p]),
    "mqct_hip_batched_profile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "mqct_hip_batched_tile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_int_p, c_int_p, c_int_p, c_int_p]),
    "mqct_hip_batched_set_tile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "mqct_hip_batched_set_tile_levels": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_int_p]),
    "mqct_hip_batched_set_step_threads": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "mqct_hip_set_contract_variant": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "mqct_hip_contract_variant": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_int_p, c_int_p, c_int_p, c_int_p]),
    "mqct_hip_batched_prof_trace": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_dbl_p, c_int_p]),
    "mqct_hip_batched_tile_levels": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_int_p]),
    "mqct_hip_batched_kernel_stats": (ctypes.c_int, [ctypes.c_void_p, c_i64_p, c_dbl_p, c_dbl_p, c_dbl_p]),
    "mqct_hip_propagate_record": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(TrajT),
                                                 ctypes.c_int, c_dbl_p, c_int_p, c_dbl_p,
                                                 ctypes.POINTER(TrajStatT)]),
    "mqct_hip_cross_sections": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_dbl_p,
                                               ctypes.c_int, c_int_p, ctypes.c_int, c_dbl_p, c_dbl_p, c_dbl_p]),
    "mqct_hip_cross_sections_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                       ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "mqct_hip_write_mtrx_uf": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int64, ctypes.c_int, c_int_p, c_int_p, c_int_p, c_int_p,
                                              c_int_p, c_dbl_p, c_dbl_p]),
    "mqct_hip_read_mtrx_uf": (ctypes.c_int, [ctypes.c_char_p, c_i64_p, c_int_p, c_int_p, c_int_p, c_int_p,
                                             c_int_p, c_dbl_p, c_dbl_p]),
    "mqct_hip_write_at_path": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, c_dbl_p]),
    "mqct_hip_read_at_path": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, c_dbl_p, c_int_p]),
    "mqct_hip_mlp_init": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64]),
    "mqct_hip_mlp_n_params": (ctypes.c_int, []),
    "mqct_hip_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, c_flt_p]),
    "mqct_hip_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, c_flt_p]),
    "mqct_hip_mlp_set_scaling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double]),
    "mqct_hip_mlp_set_data": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_dbl_p, c_flt_p]),
    "mqct_hip_mlp_train_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                               c_flt_p]),
    "mqct_hip_mlp_grad": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_flt_p, c_flt_p]),
    "mqct_hip_mlp_predict": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_dbl_p, c_flt_p]),
    "mqct_hip_surrogate_vgrid": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_dbl_p]),
    "mqct_hip_pes_points": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_dbl_p, ctypes.c_int,
                                           c_dbl_p, c_dbl_p]),
    "mqct_hip_probe_mfma": (ctypes.c_int, [ctypes.c_int, c_dbl_p, c_dbl_p, c_dbl_p]),
}

_lib = None


def load(path=LIB_PATH):
    """Load the library once; raise LibraryMissing if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LibraryMissing(
            f"{path} not found: build it with `python -m mqct_amd.build` "
            "(the HIP library is required; there is no fallback)")
    try:
        lib = ctypes.CDLL(path)
    except OSError as e:
        raise LibraryMissing(f"cannot load {path}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def missing_symbols(path=LIB_PATH):
    lib = ctypes.CDLL(path)
    return [n for n in SIGNATURES if not hasattr(lib, n)]


def last_error():
    lib = load()
    buf = ctypes.create_string_buffer(4096)
    lib.mqct_hip_last_error(buf, 4096)
    return buf.value.decode(errors="replace")


def check(rc, what=""):
    if rc != 0:
        raise MqctError(f"{what}: {last_error()} (status {rc})")


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be contiguous")
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def iptr(a):
    return ptr(a, ctypes.c_int)


def dptr(a):
    return ptr(a, ctypes.c_double)


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
