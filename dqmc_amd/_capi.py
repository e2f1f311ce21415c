"""ctypes binding of libdqmc_hip.so (declared in include/dqmc_hip.h).

The library is the product: importing this module never falls back to anything
else. If the shared library has not been built, loading raises immediately.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libdqmc_hip.so")


class DQMCError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [("L", ctypes.c_int), ("slices", ctypes.c_int), ("safe_mult", ctypes.c_int),
                ("checkerboard", ctypes.c_int), ("edrun", ctypes.c_int), ("all_checks", ctypes.c_int),
                ("delta_tau", ctypes.c_double), ("lam", ctypes.c_double), ("r", ctypes.c_double),
                ("c", ctypes.c_double), ("u", ctypes.c_double)]


class Operators(ctypes.Structure):
    _fields_ = [("plaq_sites", ctypes.POINTER(ctypes.c_int)), ("plaq", ctypes.POINTER(ctypes.c_double)),
                ("expmu", ctypes.POINTER(ctypes.c_double)), ("hop_exp", ctypes.POINTER(ctypes.c_double)),
                ("hop_exp_inv", ctypes.POINTER(ctypes.c_double)), ("neighbors", ctypes.POINTER(ctypes.c_int)),
                ("n_groups", ctypes.c_int), ("group_offsets", ctypes.POINTER(ctypes.c_int)),
                ("bond_sites", ctypes.POINTER(ctypes.c_int)), ("bond_blocks", ctypes.POINTER(ctypes.c_double))]


class Stats(ctypes.Structure):
    _fields_ = [("current_slice", ctypes.c_int), ("direction", ctypes.c_int),
                ("boson_action", ctypes.c_double), ("log_det", ctypes.c_double),
                ("drift_last", ctypes.c_double), ("drift_max", ctypes.c_double),
                ("acc_last", ctypes.c_double), ("accepted", ctypes.c_double), ("proposed", ctypes.c_double),
                ("nonreal", ctypes.c_double), ("min_ratio_gap", ctypes.c_double),
                ("uniforms_used", ctypes.c_longlong), ("device_error", ctypes.c_int),
                ("global_accepted", ctypes.c_double), ("global_proposed", ctypes.c_double)]


_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_LL = ctypes.POINTER(ctypes.c_longlong)
_i = ctypes.c_int
_d = ctypes.c_double
_ll = ctypes.c_longlong

# name -> argtypes (all return int status unless listed in _RET)
SIGNATURES = {
    "dqmc_ctx_create": [ctypes.POINTER(Params), ctypes.POINTER(Operators), _i, ctypes.POINTER(_P)],
    "dqmc_ctx_destroy": [_P],
    "dqmc_ctx_set_groups": [_P, _i, _i, _i],
    "dqmc_set_greens_route": [_P, _i],
    "dqmc_set_local_variant": [_P, _i],
    "dqmc_set_udt_route": [_P, _i],
    "dqmc_set_field": [_P, _D],
    "dqmc_get_field": [_P, _D],
    "dqmc_set_uniforms": [_P, _D, _ll],
    "dqmc_init": [_P],
    "dqmc_build_stack": [_P],
    "dqmc_propagate": [_P, _I, _I, _D],
    "dqmc_local_updates": [_P, _d, _D],
    "dqmc_sweep": [_P, _i, _d],
    "dqmc_synchronize": [_P],
    "dqmc_get_greens": [_P, _D],
    "dqmc_set_greens": [_P, _D],
    "dqmc_get_stack": [_P, _i, _D, _D, _D],
    "dqmc_get_stats": [_P, ctypes.POINTER(Stats)],
    "dqmc_reset_stats": [_P],
    "dqmc_set_proposal_log": [_P, _ll],
    "dqmc_get_proposal_log": [_P, _D, _LL],
    "dqmc_apply_B": [_P, _i, _i, _D],
    "dqmc_wrap_greens": [_P, _D, _i, _i],
    "dqmc_udt": [_P, _D, _D, _D, _D],
    "dqmc_udt_prepivoted": [_P, _D, _D, _D, _D],
    "dqmc_calculate_greens": [_P, _D, _D, _D, _D, _D, _D, _D, _D],
    "dqmc_greens_at": [_P, _i, _D, _D],
    "dqmc_detratio": [_P, _i, _D, _D, _D, _D],
    "dqmc_update_greens": [_P, _i],
    "dqmc_gemm": [_P, _i, _i, _D, _D, _D],
    "dqmc_inverse": [_P, _D, _D, _D],
    "dqmc_boson_action": [_P, _D],
    "dqmc_inverse_noalloc": [_i, _P, _P, _P, _P],
    "dqmc_global_update": [_P, _d, _I],
    "dqmc_effective_greens2greens": [_P, _D],
    "dqmc_tdgf_alloc": [_P],
    "dqmc_tdgf_free": [_P],
    "dqmc_measure_tdgfs": [_P],
    "dqmc_get_tdgf": [_P, _i, _D, _D],
    "dqmc_get_tdgf_stack": [_P, _i, _i, _D, _D, _D],
    "dqmc_tdgf_direct": [_P, _i, _D],
    "dqmc_measure_greens": [_P, _D],
    "dqmc_measure_current_greens": [_P, _D],
    "dqmc_set_etgf": [_P, _D],
    "dqmc_set_hopping_matrix": [_P, _D],
    "dqmc_measure_etpc": [_P, _D, _D],
    "dqmc_measure_etcdc": [_P, _D, _D],
    "dqmc_measure_zfpc": [_P, _D, _D],
    "dqmc_measure_zfcdc": [_P, _D, _D],
    "dqmc_measure_zfccc": [_P, _D, _D, _D],
    "dqmc_grad": [_P, _D, _D],
    "dqmc_hmc_trajectory": [_P, _D, _i, _d, _d, _I, _D],
    "dqmc_batch_create": [ctypes.POINTER(Params), ctypes.POINTER(Operators), _i, _i, ctypes.POINTER(_P)],
    "dqmc_batch_destroy": [_P],
    "dqmc_batch_set_field": [_P, _i, _D],
    "dqmc_batch_get_field": [_P, _i, _D],
    "dqmc_batch_set_uniforms": [_P, _i, _D, _ll],
    "dqmc_batch_init": [_P],
    "dqmc_batch_sweep": [_P, _i, _D],
    "dqmc_batch_synchronize": [_P],
    "dqmc_batch_get_greens": [_P, _i, _D],
    "dqmc_batch_get_stats": [_P, _i, ctypes.POINTER(Stats)],
    "dqmc_batch_use_graph": [_P, _i],
    "dqmc_batch_set_merged": [_P, _i],
    "dqmc_use_graph": [_P, _i],
    "dqmc_set_qr_panel": [_P, _i],
    "dqmc_batch_set_qr_panel": [_P, _i],
    "dqmc_batch_autotune": [_P, _D, _i, _D],
    "dqmc_autotune": [_P, ctypes.c_double, _i, _D],
    "dqmc_batch_set_groups": [_P, _i, _i, _i],
    "dqmc_batch_set_local_variant": [_P, _i],
    "dqmc_profile": [_P, ctypes.c_uint],
    "dqmc_profile_read": [_P, _D, _LL, _i],
    "dqmc_batch_profile": [_P, ctypes.c_uint],
    "dqmc_batch_profile_read": [_P, _D, _LL, _i],
    "dqmc_last_error": [],
}
_RET = {"dqmc_last_error": ctypes.c_char_p}

# kernel classes of dqmc_profile (include/dqmc_hip.h)
K_APPLY_B, K_GEMM, K_QRP, K_QFORM, K_INVERSE, K_LOCAL, K_OTHER = range(7)
K_NCLASS = 7
K_NAMES = ["apply_B", "gemm", "qrp", "qform", "inverse", "local_updates", "other"]

_lib = None


def load(path=LIB_PATH):
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise DQMCError("libdqmc_hip.so not built at %s (run `python -m dqmc_amd.build`)" % path)
    lib = ctypes.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RET.get(name, ctypes.c_int)
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = load().dqmc_last_error().decode(errors="replace")
        raise DQMCError("dqmc error %d: %s" % (rc, msg))


def dptr(a):
    """double* of a C-contiguous float64/complex128 numpy array (None -> NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(_D)


def iptr(a):
    return a.ctypes.data_as(_I)
