"""ctypes binding of libmckl_hip.so (the C ABI declared in include/mckl_hip.h).

The library is required: importing this module raises if it is missing, so
no caller can silently fall back to a CPU path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmckl_hip.so")

OK = 0
EINVAL = -22

PHILOX4X32 = 1
THREEFRY4X64 = 2
ARS = 3

U01_CC, U01_CO, U01_OC, U01_OO, U01_CANONICAL = 0, 1, 2, 3, 4

MULTINOMIAL, STRATIFIED, SYSTEMATIC = 0, 1, 2
RESIDUAL, RESIDUAL_STRATIFIED, RESIDUAL_SYSTEMATIC = 3, 4, 5

MODEL_LG, MODEL_SV = 1, 2

ROW_MAJOR, COL_MAJOR = 101, 102


class Engine(ctypes.Structure):
    """mckl_hip_engine: key, counter and consumed-word index."""

    _fields_ = [
        ("ctr", ctypes.c_uint64 * 4),
        ("key", ctypes.c_uint64 * 4),
        ("index", ctypes.c_uint32),
        ("kind", ctypes.c_int32),
        ("rbits", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class Tune(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int32), ("waves_per_eu", ctypes.c_int32),
                ("split_tiles", ctypes.c_int32),
                ("small_block", ctypes.c_int32)]


if not os.path.exists(LIB_PATH):
    raise ImportError(
        "libmckl_hip.so is not built (expected at %s); run "
        "`python -m mckl_amd.build`" % LIB_PATH)

# torch ships its own libamdhip64.so.7 (the same soname as ROCm's): import it
# first so that this library binds to the HIP runtime torch already uses;
# loading ROCm's copy first leaves torch without a device.
import torch  # noqa: E402,F401

lib = ctypes.CDLL(LIB_PATH)

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_u64 = ctypes.c_uint64
_int = ctypes.c_int
_dbl = ctypes.c_double
_flt = ctypes.c_float
_EP = ctypes.POINTER(Engine)
_U64P = ctypes.POINTER(ctypes.c_uint64)
_U32P = ctypes.POINTER(ctypes.c_uint32)

SIGNATURES = {
    "mckl_hip_last_error": (ctypes.c_char_p, []),
    "mckl_hip_engine_words": (_int, [_EP]),
    "mckl_hip_engine_init": (_int, [_EP, _int, _int, _U64P]),
    "mckl_hip_engine_seed": (_int, [_EP, _u64]),
    "mckl_hip_engine_set_ctr": (_int, [_EP, _U64P]),
    "mckl_hip_engine_discard": (_int, [_EP, _u64]),
    "mckl_hip_engine_next": (_int, [_EP, _U64P]),
    "mckl_hip_engine_block": (_int, [_EP, _vp]),
    "mckl_hip_philox4x32_kernel": (_int, [_U64P, _sz, _vp, _U32P, _vp]),
    "mckl_hip_threefry4x64_kernel": (_int, [_U64P, _sz, _vp, _U64P, _vp]),
    "mckl_hip_ars_kernel": (_int, [_U64P, _sz, _vp, _U32P, _vp]),
    "mckl_hip_rand": (_int, [_EP, _sz, _vp, _vp]),
    "mckl_hip_rand_uniform_bits": (_int, [_EP, _int, _sz, _vp, _vp]),
    "mckl_hip_rand_u01_f32": (_int, [_EP, _int, _sz, _vp, _vp]),
    "mckl_hip_rand_u01_f64": (_int, [_EP, _int, _sz, _vp, _vp]),
    "mckl_hip_rand_normal_f32": (_int, [_EP, _sz, _flt, _flt, _vp, _vp]),
    "mckl_hip_rand_normal_f64": (_int, [_EP, _sz, _dbl, _dbl, _vp, _vp]),
    "mckl_hip_rand_uniform_real_f32": (_int, [_EP, _sz, _flt, _flt, _vp, _vp]),
    "mckl_hip_rand_uniform_real_f64": (_int, [_EP, _sz, _dbl, _dbl, _vp, _vp]),
    "mckl_hip_rand_u01_f64_streams": (_int, [_EP, _sz, _int, _sz, _vp, _vp]),
    "mckl_hip_rand_normal_f32_streams": (
        _int, [_EP, _sz, _sz, _flt, _flt, _vp, _vp]),
    "mckl_hip_rand_normal_f64_streams": (
        _int, [_EP, _sz, _sz, _dbl, _dbl, _vp, _vp]),
    "mckl_hip_weight_set_equal": (_int, [_sz, _vp, _vp, _vp]),
    "mckl_hip_weight_set_log": (_int, [_sz, _vp, _vp, _vp, _vp]),
    "mckl_hip_weight_add_log": (_int, [_sz, _vp, _vp, _vp, _vp]),
    "mckl_hip_u01_rand_sequence": (_int, [_int, _EP, _sz, _vp, _vp]),
    "mckl_hip_resample_trans_u01_rep": (_int, [_sz, _sz, _vp, _vp, _vp, _vp]),
    "mckl_hip_resample_trans_rep_index": (_int, [_sz, _sz, _vp, _vp, _vp]),
    "mckl_hip_resample": (_int, [_int, _sz, _sz, _EP, _vp, _vp, _vp, _vp]),
    "mckl_hip_resample_noalloc": (
        _int, [_int, _sz, _sz, _EP, _vp, _vp, _vp, _vp]),
    "mckl_hip_select_f32": (_int, [_sz, _sz, _vp, _vp, _vp, _vp]),
    "mckl_hip_select_f64": (_int, [_sz, _sz, _vp, _vp, _vp, _vp]),
    "mckl_hip_resample_index_create": (_int, [ctypes.POINTER(_vp)]),
    "mckl_hip_resample_index_destroy": (_int, [_vp]),
    "mckl_hip_resample_index_num_iter": (_int, [_vp, ctypes.POINTER(_sz)]),
    "mckl_hip_resample_index_size": (_int, [_vp, _sz, ctypes.POINTER(_sz)]),
    "mckl_hip_resample_index_reset": (_int, [_vp]),
    "mckl_hip_resample_index_clear": (_int, [_vp]),
    "mckl_hip_resample_index_push_back": (_int, [_vp, _sz, _vp, _vp]),
    "mckl_hip_resample_index_index": (_int, [_vp, _sz, _vp, _sz, _sz, _vp, _vp]),
    "mckl_hip_resample_index_matrix_shape": (
        _int, [_vp, _sz, _sz, ctypes.POINTER(_sz), ctypes.POINTER(_sz)]),
    "mckl_hip_resample_index_read_matrix": (
        _int, [_vp, _int, _sz, _sz, _vp, _vp]),
    "mckl_hip_smc_create": (
        _int, [_int, ctypes.POINTER(_dbl), _sz, ctypes.POINTER(_dbl), _sz,
               _u64, _dbl, _int, ctypes.POINTER(_vp)]),
    "mckl_hip_smc_run": (_int, [_vp, _vp]),
    "mckl_hip_smc_result": (
        _int, [_vp, ctypes.POINTER(_dbl), ctypes.POINTER(_dbl),
               ctypes.POINTER(ctypes.c_int32)]),
    "mckl_hip_smc_destroy": (_int, [_vp]),
    "mckl_hip_smc_estimate_moments": (_int, [_vp, _int]),
    "mckl_hip_smc_estimates": (_int, [_vp, ctypes.POINTER(_dbl)]),
    "mckl_hip_estimate_f64": (_int, [_sz, _sz, _int, _vp, _vp, _vp, _vp]),
    "mckl_hip_estimate_f32": (_int, [_sz, _sz, _int, _vp, _vp, _vp, _vp]),
    "mckl_hip_smc_batch_create": (
        _int, [_int, _sz, _U32P, ctypes.POINTER(_dbl), _sz,
               ctypes.POINTER(_dbl), _U64P, _dbl, _int, ctypes.POINTER(_vp)]),
    "mckl_hip_smc_batch_set_tune": (_int, [_vp, ctypes.POINTER(Tune)]),
    "mckl_hip_smc_batch_run": (_int, [_vp, _int, _vp]),
    "mckl_hip_smc_batch_autotune": (
        _int, [_vp, ctypes.c_char_p, ctypes.POINTER(Tune),
               ctypes.POINTER(_dbl), _vp]),
    "mckl_hip_smc_batch_result": (_int, [_vp, ctypes.POINTER(_dbl)]),
    "mckl_hip_smc_batch_destroy": (_int, [_vp]),
    "mckl_hip_smc_batch_set_params": (
        _int, [_vp, ctypes.POINTER(_dbl), _U64P]),
    "mckl_hip_seed_keys": (
        _int, [_int, _u64, _u64, _u64, _int, _sz, _vp, _vp]),
    "mckl_hip_rngset_rand": (_int, [_int, _sz, _vp, _sz, _vp, _vp]),
    "mckl_hip_rngset_u01_f32": (_int, [_int, _int, _sz, _vp, _sz, _vp, _vp]),
    "mckl_hip_rngset_u01_f64": (_int, [_int, _int, _sz, _vp, _sz, _vp, _vp]),
    "mckl_hip_smc_grad_create": (
        _int, [_int, _sz, _sz, _sz, ctypes.POINTER(_dbl), _int, _U64P, _dbl,
               _int, _sz, _int, ctypes.POINTER(_vp)]),
    "mckl_hip_smc_grad_run": (
        _int, [_vp, ctypes.POINTER(_dbl), ctypes.POINTER(_dbl),
               ctypes.POINTER(_dbl), _vp]),
    "mckl_hip_smc_grad_replay": (
        _int, [_vp, ctypes.POINTER(_dbl), ctypes.POINTER(_dbl), _vp]),
    "mckl_hip_smc_grad_destroy": (_int, [_vp]),
    "mckl_hip_smc_loglik_grad": (
        _int, [_int, _sz, ctypes.POINTER(_dbl), _sz, _sz, ctypes.POINTER(_dbl),
               _int, _U64P, _dbl, _int, _sz, ctypes.POINTER(_dbl),
               ctypes.POINTER(_dbl), _vp]),
}

MISSING = []
for _name, (_res, _args) in SIGNATURES.items():
    try:
        _f = getattr(lib, _name)
    except AttributeError:
        MISSING.append(_name)
        continue
    _f.restype = _res
    _f.argtypes = _args


class MCKLError(RuntimeError):
    pass


def check(rc):
    if rc != OK:
        msg = lib.mckl_hip_last_error()
        raise MCKLError("mckl_hip error %d: %s" % (rc, msg.decode() if msg else ""))
    return rc
