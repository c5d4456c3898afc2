"""Loads libnecstar_hip.so and declares the C-ABI of include/necstar_hip.h."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnecstar_hip.so")

NX_OK, NX_E_VALUE, NX_E_TYPE, NX_E_DEVICE, NX_E_INTERNAL = range(5)

_P = C.c_void_p
_I32, _I64, _U64, _D = C.c_int32, C.c_int64, C.c_uint64, C.c_double
_I32P, _I64P, _U64P = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_uint64)
_U8P, _U32P, _DP = C.POINTER(C.c_uint8), C.POINTER(C.c_uint32), C.POINTER(C.c_double)

SIGNATURES = {
    "nx_last_error": (C.c_char_p, []),
    "nx_init": (C.c_int, [C.c_int]),
    "nx_set_stream": (C.c_int, [_P, C.c_int]),
    "nx_random_clifford": (C.c_int, [_I32, _U8P, _I64, _I32P, _I64P]),
    "nx_state_compile": (C.c_int, [C.c_int, _I64, _I32P, C.POINTER(_P)]),
    "nx_state_clone": (C.c_int, [_P, C.POINTER(_P)]),
    "nx_state_free": (None, [_P]),
    "nx_state_num_qubits": (C.c_int, [_P, _I32P]),
    "nx_state_stabilizer_rank": (C.c_int, [_P, _I64P]),
    "nx_state_apply_gates": (C.c_int, [_P, _I64, _I32P]),
    "nx_state_norm": (C.c_int, [_P, _DP]),
    "nx_state_norm_noalloc": (C.c_int, [_P, _DP]),
    "nx_state_norm_async": (C.c_int, [_P, _DP, _P]),
    "nx_state_norm_squared": (C.c_int, [_P, _DP]),
    "nx_state_exp_value": (C.c_int, [_P, C.c_int, _I32, _I32P, _I32P, _DP]),
    "nx_state_inner_product": (C.c_int, [_P, _P, _DP, _DP]),
    "nx_state_amplitudes": (C.c_int, [_P, _I64, _U64P, _I32, _DP, _U8P, _I32P]),
    "nx_state_amplitudes_device": (C.c_int, [_P, _I64, _U64P, _I32, _DP]),
    "nx_state_amplitudes_async": (C.c_int, [_P, _I64, _U64P, _I32, _DP, _P]),
    "nx_state_statevector": (C.c_int, [_P, _DP]),
    "nx_state_project": (C.c_int, [_P, _I32, _I32, _I32]),
    "nx_state_discard": (C.c_int, [_P, _I32]),
    "nx_state_measure": (C.c_int, [_P, _I32, _I32P, _U8P, _U8P]),
    "nx_state_sample": (C.c_int, [_P, _I32, _I32P, _U64, _U8P, _I64, _U64P, _U64P, _I64P]),
    "nx_state_sample_exact": (C.c_int, [_P, _I32, _I32P, _U64, _U8P, _I64, _U64P, _U64P, _I64P]),
    "nx_state_export_term": (C.c_int, [_P, _I64, _U8P, _U8P, _U8P, _U8P, _U8P, _U8P, _U8P]),
    "nx_state_get_scalar_coefficients": (C.c_int, [_P, _U8P, _U8P, _I64P]),
    "nx_state_get_coefficients": (C.c_int, [_P, _DP]),
    "nx_state_set_coefficients": (C.c_int, [_P, _DP]),
    "nx_state_global_factor": (C.c_int, [_P, _DP, _DP]),
    "nx_state_pair_table": (C.c_int, [_P, _U32P]),
    "nx_state_cross_pair_table": (C.c_int, [_P, _P, _U32P]),
    "nx_state_grad_exp_values": (C.c_int, [_P, _I32, _I32P, _DP, _DP, _DP]),
    "nx_state_grad_exp_values_device": (C.c_int, [_P, _I32, _I32P, _DP, _DP, _DP, _DP, _P]),
    "nx_state_exp_values": (C.c_int, [_P, _I32, _I32P, _DP]),
    "nx_profile_enable": (C.c_int, [C.c_int]),
    "nx_profile_only": (C.c_int, [C.c_char_p]),
    "nx_profile_get": (C.c_int, [C.c_char_p, _DP, _I64P]),
    "nx_state_drop_cache": (C.c_int, [_P]),
    "nx_state_projection_norms": (C.c_int, [_P, _I32, _I32P, _I32P, _I32, _I32P, _DP]),
    "nx_state_uniform_prefix": (C.c_int, [_P, _I32, _I32P, _I32P]),
    "nx_state_uniform_prefix_given": (C.c_int, [_P, _I32, _I32P, _I32P, _I32, _I32P, _I32P]),
    "nx_batch_create": (C.c_int, [_I32, C.POINTER(_P), _I32P, _I32P, _I32, C.POINTER(_P)]),
    "nx_batch_num_outputs": (C.c_int, [_P, _I64P]),
    "nx_batch_shape": (C.c_int, [_P, _I32P, _I64P, _I32P]),
    "nx_batch_run": (C.c_int, [_P, _DP]),
    "nx_batch_free": (None, [_P]),
    "nx_state_completion_norms": (C.c_int, [_P, _U64P, _I32, _I32P, _DP, _DP]),
    "nx_batch_sample_exact": (C.c_int, [_I32, C.POINTER(_P), _I32P, _I32P, _U64P, _U8P, _I64, _U64P, _U64P,
                                        _I64P, _I64P]),
}

_lib = None


def lib():
    """The loaded C-ABI library.  Fails loudly when the HIP extension is
    missing: there is no CPU fallback behind the necstar module."""
    global _lib
    if _lib is None:
        # PyTorch-ROCm carries its own HIP runtime.  When torch is installed
        # it is imported first, so that this library binds to that runtime
        # instead of loading a second one (with two runtimes in one process,
        # a torch imported afterwards finds no GPU).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(hipcc --offload-arch=gfx950); necstar has no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def loaded():
    """The library if it is already loaded, else None (for destructors, which
    may run during interpreter shutdown and must not load anything)."""
    return _lib


def check(code: int) -> None:
    if code == NX_OK:
        return
    msg = lib().nx_last_error().decode()
    if code == NX_E_VALUE:
        raise ValueError(msg)
    if code == NX_E_TYPE:
        raise TypeError(msg)
    raise RuntimeError(msg)
