"""ctypes binding of libofdft_hip.so (the C ABI in include/ofdft_hip.h).

The library is loaded from ``dftpy_amd/lib`` after ``torch`` so both share
one HIP runtime; there is no fallback: a missing library raises at import.
"""
import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime the library binds to)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libofdft_hip.so")

FP64, FP32 = 0, 1
CALC_E, CALC_V = 1, 2
TERM_TF, TERM_VW, TERM_WT, TERM_XC, TERM_HARTREE, TERM_PP, TERM_LMGP = 1, 2, 4, 8, 16, 32, 64
TERMS_ALL = 63
E_TF, E_VW, E_NL, E_XC, E_H, E_PP, E_SUM, E_NELEC, E_COUNT = range(9)

if not os.path.exists(LIB_PATH):
    raise ImportError("libofdft_hip.so not built (%s); run `python -m dftpy_amd.build`" % LIB_PATH)

lib = ctypes.CDLL(LIB_PATH)

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)

SIGNATURES = {
    "ofdft_version": (_I, []),
    "ofdft_last_error": (ctypes.c_char_p, []),
    "ofdft_plan_create": (_I, [_IP, _DP, _I, _I, _I, ctypes.POINTER(_P)]),
    "ofdft_plan_destroy": (_I, [_P]),
    "ofdft_plan_set_lattice": (_I, [_P, _DP]),
    "ofdft_plan_workspace_bytes": (ctypes.c_size_t, [_P]),
    "ofdft_plan_set_fft_kernels": (_I, [_P, _I]),
    "ofdft_plan_autotune": (_I, [_P, _IP, ctypes.POINTER(ctypes.c_float), _P]),
    "ofdft_plan_set_variants": (_I, [_P, _IP]),
    "ofdft_fft_r2c": (_I, [_P, _P, _P, _I, _P]),
    "ofdft_fft_c2r": (_I, [_P, _P, _P, _I, _P]),
    "ofdft_fft_c2c": (_I, [_P, _P, _P, _I, _I, _P]),
    "ofdft_set_kedf": (_I, [_P, _D, _D, _D, _D, _D, _D]),
    "ofdft_set_terms": (_I, [_P, ctypes.c_uint]),
    "ofdft_set_vloc": (_I, [_P, _P]),
    "ofdft_set_core": (_I, [_P, _P]),
    "ofdft_set_eion": (_I, [_P, _D]),
    "ofdft_mgp_kernel_table": (_I, [_I, _D, _I, _I, _DP, _DP, _DP]),
    "ofdft_set_lmgp": (_I, [_P, _I, _D, _DP, _DP, _D]),
    "ofdft_lmgp_reset": (_I, [_P]),
    "ofdft_eval": (_I, [_P, _P, ctypes.c_uint, _P, _P, _P]),
    "ofdft_eval_fields": (_I, [_P, _P, ctypes.c_uint, _P, _P, _P, _P, _P]),
    "ofdft_dot": (_I, [_P, _P, _P, _P, _P]),
    "ofdft_bench_pass": (_I, [_P, _I, _I, _I, _P, _DP, _P]),
    "ofdft_atomic_density": (_I, [_P, _I, _I, _IP, ctypes.POINTER(_DP), ctypes.POINTER(_DP), _DP, _DP, _DP, _DP, _I,
                                  _DP, _IP, _P, _P]),
    "ofdft_atomic_density_noalloc": (_I, [_P, _I, _I, _IP, ctypes.POINTER(_DP), ctypes.POINTER(_DP), _DP, _DP, _DP,
                                          _DP, _I, _DP, _IP, _P, _P]),
    "ofdft_clamp_normalize": (_I, [_P, _P, _I, _D, _D, _P]),
    "ofdft_local_pp": (_I, [_P, _I, _DP, _DP, _IP, ctypes.POINTER(_DP), _DP, _DP, _DP, _I, _DP, _DP, _IP, _I, _I,
                            _P, _P]),
    "ofdft_ewald": (_I, [_P, _I, _DP, _DP, _DP, _I, _I, _D, _DP, _P]),
    "ofdft_optimize": (_I, [_P, _P, _P, _I, _I, _DP, _DP, _DP, _I, _P]),
    "ofdft_optimize_tn_graph": (_I, [_P, _P, _P, _I, _DP, _I, _I, _DP, _DP, _I, _P]),
    "ofdft_set_tn_precond": (_I, [_P, _I, _D, _DP]),
    "ofdft_set_tn_precond_uniform": (_I, [_P, _I, _D, _D, _DP, _P]),
    "ofdft_get_tn_precond": (_I, [_P, _P, _I, _P]),
    "ofdft_screened_superposition": (_I, [_P, _I, _DP, _DP, _IP, ctypes.POINTER(_DP), _DP, _DP, _DP, _DP, _DP, _DP,
                                          _DP, _I, _DP, _DP, _IP, _I, _I, _P, _P]),
    "ofdft_plan_reserve_lbfgs": (_I, [_P, _I]),
    "ofdft_local_pp_forces": (_I, [_P, _P, _I, _DP, _DP, _IP, ctypes.POINTER(_DP), _DP, _DP, _DP, _I, _DP, _DP, _IP,
                                   _DP, _P]),
    "ofdft_local_pp_stress": (_I, [_P, _P, _I, _DP, _DP, _IP, ctypes.POINTER(_DP), _DP, _DP, _DP, _I, _DP, _DP, _IP,
                                   _D, _DP, _P]),
    "ofdft_ewald_forces_stress": (_I, [_P, _I, _DP, _DP, _DP, _D, _DP, _DP, _P]),
    "ofdft_local_pp_forces_pme": (_I, [_P, _P, _I, _DP, _DP, _IP, ctypes.POINTER(_DP), _DP, _DP, _DP, _I, _DP, _DP, _IP,
                                       _I, _DP, _P]),
    "ofdft_local_pp_stress_pme": (_I, [_P, _P, _I, _DP, _DP, _IP, ctypes.POINTER(_DP), _DP, _DP, _DP, _I, _DP, _DP, _IP,
                                       _I, _D, _DP, _P]),
    "ofdft_ewald_forces_stress_pme": (_I, [_P, _I, _DP, _DP, _DP, _D, _I, _DP, _DP, _P]),
    "ofdft_functional_stress": (_I, [_P, _P, _DP, _DP, _P]),
    "ofdft_dcsrch": (_I, [_DP, _D, _D, _D, _D, _D, _D, _D, _IP, _DP]),
    "ofdft_cn_step": (_I, [_P, _P, _P, _D, _D, _D, _I, _P, _DP, _P]),
    "ofdft_cn_solver_workspace_bytes": (ctypes.c_size_t, [_P, _I, _I]),
    "ofdft_cn_step_solver": (_I, [_P, _I, _I, _P, _P, _D, _D, _D, _I, _P, _DP, _P, ctypes.c_size_t, _P]),
    "ofdft_cn_batch_workspace_bytes": (ctypes.c_size_t, [_P, _I]),
    "ofdft_cn_step_batch": (_I, [_P, _I, _P, _P, _D, _D, _D, _I, _P, _DP, _P, ctypes.c_size_t, _P]),
    "ofdft_cnn_conv_in": (_I, [_IP, _I, _P, _P, _P, _P, _P]),
    "ofdft_cnn_conv_mid": (_I, [_IP, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "ofdft_cnn_conv_out": (_I, [_IP, _I, _P, _P, _P, _P, _P, _P]),
    "ofdft_cnn_out_backward": (_I, [_IP, _I, _P, _P, _P, _P, _P, _P]),
    "ofdft_cnn_wgrad_partials": (ctypes.c_size_t, [_I, _I]),
    "ofdft_cnn_wgrad": (_I, [_IP, _I, _I, _P, _P, _P, _I, _P, _P]),
    "ofdft_cnn_prep_weights": (_I, [_P, _P, _P, _P]),
    "ofdft_cnn_loss": (_I, [_P, _P, _P, _P, _P, _P]),
    "ofdft_cnn_loss_spectral": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "ofdft_cnn_adam": (_I, [ctypes.c_longlong, _P, _P, _P, _P, _D, _D, _D, _D, _I, _P]),
    "ofdft_cnn_features": (_I, [_P, _P, _P, _D, _D, _D, _P, _P]),
    "ofdft_cnn_target": (_I, [_P, _P, _P, _D, _P, _P]),
    "ofdft_cnn_apply": (_I, [_P, _P, _P, _D, _D, _P, _P]),
}

for _name, (_res, _args) in SIGNATURES.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


class OFDFTError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise OFDFTError(lib.ofdft_last_error().decode())


def dptr(arr):
    """ctypes double* of a contiguous float64 numpy array."""
    return arr.ctypes.data_as(_DP)


def iptr(arr):
    return arr.ctypes.data_as(_IP)


def stream():
    """hipStream_t of torch's current stream (kernels order with torch work)."""
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def require_gpu():
    if not torch.cuda.is_available():
        raise OFDFTError("dftpy_amd needs a HIP GPU (none visible); there is no CPU path")
