"""ctypes binding of libqclhip.so (the C-ABI declared in include/qclhip.h).

The library is required: importing this module on a machine without the
built library raises immediately; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as ct
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "lib", "libqclhip.so"))

OK = 0
ERR_ARGUMENT = -1
ERR_DIMENSION = -2
ERR_DOMAIN = -3
ERR_RUNTIME = -4
ERR_UNSUPPORTED = -5

KERNEL_DENSE = 0
KERNEL_MOMENTUM_INDEPENDENT = 1
KERNEL_SEPARABLE_IFR = 2
KERNEL_LOW_RANK = 3


class QCLHipError(RuntimeError):
    """Base class; the subclasses mirror the Julia exception types."""


class SingularException(QCLHipError):
    def __init__(self, block: int, msg: str = ""):
        super().__init__(f"singular Dyson block {block} {msg}".strip())
        self.block = block


class ArgumentError(QCLHipError):
    pass


class DimensionMismatch(QCLHipError):
    pass


class DomainError(QCLHipError):
    pass


class UnsupportedError(QCLHipError):
    pass


_vp = ct.c_void_p
_dp = ct.POINTER(ct.c_double)
_i64 = ct.c_int64

_SIGNATURES = {
    "qclhip_ctx_create": [ct.c_int, _vp, ct.POINTER(_vp)],
    "qclhip_ctx_destroy": [_vp],
    "qclhip_ctx_synchronize": [_vp],
    "qclhip_ctx_last_error": [_vp],
    "qclhip_ctx_set_stream": [_vp, _vp],
    "qclhip_ctx_defer_results": [_vp, _vp, ct.c_int],
    "qclhip_ctx_result_slot": [_vp],
    "qclhip_graph_begin": [_vp],
    "qclhip_graph_end": [_vp, ct.POINTER(_vp)],
    "qclhip_graph_launch": [_vp, _vp],
    "qclhip_graph_destroy": [_vp],
    "qclhip_max_basis": [],
    "qclhip_grid_create": [_vp, ct.c_int, ct.c_int, ct.c_int, ct.c_int, _dp, _dp, _dp, ct.POINTER(_vp)],
    "qclhip_grid_destroy": [_vp],
    "qclhip_grid_set_precision": [_vp, _vp, ct.c_int, ct.c_int, ct.c_double],
    "qclhip_grid_set_engine": [_vp, _vp, ct.c_int],
    "qclhip_grid_refine_fallbacks": [_vp, _vp, ct.POINTER(ct.c_int64)],
    "qclhip_grid_autotune": [_vp, _vp, _dp],
    "qclhip_retarded_green": [_vp, _vp, _vp, _vp, ct.c_double, _vp, _vp],
    "qclhip_dyson_keldysh": [_vp, _vp, _vp, _vp, _vp, _vp, ct.c_double, _vp, _vp, _vp, _vp, _vp],
    "qclhip_dyson_keldysh_strided": [_vp, _vp, ct.c_int64, _vp, _vp, _vp, _vp, ct.c_double, _vp, _vp, _vp, _vp,
                                     _vp],
    "qclhip_current_green": [_vp, _vp, _vp, ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp), ct.POINTER(_vp),
                             ct.c_double, _vp, _vp, _vp, _vp, _vp, _dp],
    "qclhip_current_green_totals": [_vp, _vp, _vp, ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp), ct.POINTER(_vp),
                                    ct.c_double, _vp, _vp, _vp, _vp, _vp, _dp, _vp, _vp, _vp],
    "qclhip_current_green_screened": [_vp, _vp, _vp, ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp), ct.POINTER(_vp),
                                      ct.c_double, _vp, _vp, _vp, _vp, _vp, _dp, _vp, _vp, _vp, _vp],
    "qclhip_current_green_noalloc": [_vp, _vp, _vp, ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp), ct.POINTER(_vp),
                                     ct.c_double, _vp, _vp, _vp, _vp, _vp, _dp],
    "qclhip_current_green_scalar_lesser": [_vp, _vp, _vp, ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp),
                                           ct.POINTER(_vp), ct.c_double, _vp, _vp, _vp, _vp, _vp, _dp],
    "qclhip_number_functional": [_vp, _vp, _vp, _dp],
    "qclhip_spectral": [_vp, _vp, _vp, _vp],
    "qclhip_green_state": [_vp, _vp, _vp, _vp, _vp, _vp],
    "qclhip_normalize_pair": [_vp, _i64, _vp, _vp, ct.c_double, ct.c_double],
    "qclhip_shift_create": [_vp, _vp, ct.c_double, ct.POINTER(_vp)],
    "qclhip_shift_destroy": [_vp],
    "qclhip_apply_shift": [_vp, _vp, _i64, _vp, _vp],
    "qclhip_kernel_create": [_vp, _vp, ct.c_int, ct.c_int, _dp, _dp, _dp, ct.POINTER(_vp)],
    "qclhip_kernel_destroy": [_vp],
    "qclhip_contract": [_vp, _vp, _vp, ct.c_double, ct.c_double, ct.c_double, _vp],
    "qclhip_contract_compact": [_vp, _vp, _vp, ct.c_double, ct.c_double, _vp],
    "qclhip_broadcast_momenta": [_vp, _vp, _vp, _vp],
    "qclhip_lo_contraction": [_vp, _vp, _vp, _vp, ct.c_double, ct.c_double, _vp, _vp, _vp, _vp],
    "qclhip_hilbert_create": [_vp, _vp, _i64, ct.c_int, ct.POINTER(_vp)],
    "qclhip_hilbert_destroy": [_vp],
    "qclhip_hilbert_set_path": [_vp, _vp, ct.c_int],
    "qclhip_hilbert_set_multiplicity": [_vp, _vp, ct.c_int64],
    "qclhip_hilbert_get_path": [_vp, _vp, ct.POINTER(ct.c_int)],
    "qclhip_fft_hilbert": [_vp, _vp, _vp, _vp, _dp],
    "qclhip_retarded_self_energy": [_vp, _vp, _vp, _vp, _vp, _vp, _dp],
    "qclhip_embedding_component": [_vp, _vp, _vp, _vp, _vp, _vp],
    "qclhip_sum_fields": [_vp, _i64, ct.c_int, ct.POINTER(_vp), _vp],
    "qclhip_mix": [_vp, _i64, _vp, _vp, ct.c_double],
    "qclhip_mix_fields": [_vp, _i64, ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp), ct.c_double],
    "qclhip_boundary_flux": [_vp, _vp, _vp, _vp, _vp, _vp, _dp],
    "qclhip_state_populations": [_vp, _vp, _vp, ct.c_double, _dp],
    "qclhip_seed_weights": [_vp, _vp, _vp, _dp],
    "qclhip_seed_lesser": [_vp, _vp, _vp, ct.c_double, ct.c_double, _vp],
    "qclhip_current_green_adjoint": [_vp, _vp, ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp), ct.c_double, _dp,
                                     _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "qclhip_boundary_flux_adjoint": [_vp, _vp, ct.c_double, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "qclhip_shift_transpose": [_vp, _vp, _vp, _vp, ct.c_double, ct.c_int, _vp],
    "qclhip_kernel_adjoint": [_vp, _vp, ct.POINTER(_vp)],
    "qclhip_retarded_self_energy_adjoint": [_vp, _vp, _vp, _vp, _vp],
    "qclhip_embedding_adjoint": [_vp, _vp, _vp, _vp, _vp, _vp],
    "qclhip_antihermitian_part": [_vp, _vp, _vp],
    "qclhip_residual_suite": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp),
                              ct.POINTER(_vp), ct.c_int, ct.POINTER(_vp), _dp],
    "qclhip_selfenergy_residual": [_vp, ct.c_int64, ct.POINTER(_vp), ct.POINTER(_vp), _dp],
    "qclhip_electron_density": [_vp, _vp, _vp, ct.c_int, _vp, ct.c_double, _vp],
    "qclhip_periodic_poisson": [_vp, ct.c_int, _vp, _vp, _vp, ct.c_double, _vp, _vp, _dp],
    "qclhip_project_hamiltonians": [_vp, ct.c_int, ct.c_int, ct.c_int, _vp, _vp, _vp, _vp, _vp, _vp],
    "qclhip_mix_potential": [_vp, ct.c_int, _vp, _vp, ct.c_double, _dp],
    "qclhip_anderson_append": [_vp, ct.c_int64, ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp), ct.POINTER(_vp),
                               ct.POINTER(_vp), ct.c_int, ct.POINTER(_vp), _dp],
    "qclhip_anderson_combine": [_vp, ct.c_int64, ct.c_int, ct.c_int, _dp, ct.c_double, ct.POINTER(_vp),
                                ct.POINTER(_vp), ct.POINTER(_vp)],
    "qclhip_anderson_guard": [_vp, _vp, ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp), ct.POINTER(_vp),
                              ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int)],
    "qclhip_mix_fields_change": [_vp, ct.c_int64, ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp), ct.c_double, _dp],
    "qclhip_fill_zero": [_vp, ct.c_int64, _vp],
    "qclhip_mix_fields_residual": [_vp, ct.c_int64, ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp), ct.c_double,
                                   ct.POINTER(ct.c_int), _dp],
    "qclhip_mix_fields_residual_totals": [_vp, ct.c_int64, ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp), ct.c_double,
                                          ct.POINTER(ct.c_int), ct.c_int, ct.POINTER(ct.c_int),
                                          ct.POINTER(ct.c_int), ct.POINTER(_vp), _dp],
    "qclhip_mix_fields_residual_accumulate": [_vp, ct.c_int64, ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp),
                                              ct.c_double, ct.POINTER(ct.c_int), ct.c_int, ct.POINTER(ct.c_int),
                                              ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), ct.POINTER(_vp), _dp],
    "qclhip_mix_fields_residual_compact": [_vp, _vp, ct.c_int64, ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp),
                                           ct.c_double, ct.POINTER(ct.c_int), ct.c_int, ct.POINTER(ct.c_int),
                                           ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), ct.POINTER(_vp),
                                           ct.POINTER(ct.c_int), _dp],
    "qclhip_residual_state": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp),
                              ct.POINTER(_vp), _dp],
    "qclhip_residual_keldysh": [_vp, _vp, _vp, _vp, ct.c_int, ct.POINTER(_vp), _dp],
    "qclhip_residual_suite_prescreened": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ct.c_int, ct.POINTER(_vp),
                                          ct.POINTER(_vp), ct.POINTER(_vp), ct.c_int, ct.POINTER(_vp), _vp, _dp],
    "qclhip_residual_state_prescreened": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ct.c_int, ct.POINTER(_vp),
                                          ct.POINTER(_vp), ct.POINTER(_vp), _vp, _dp],
    "qclhip_batch_create": [_vp, ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp)],
    "qclhip_grid_tuning": [_vp, _vp, _dp, ct.POINTER(ct.c_int)],
    "qclhip_grid_set_keldysh_split": [_vp, _vp, ct.c_int],
    "qclhip_kernel_create_dense_device": [_vp, _vp, _vp, ct.POINTER(_vp)],
    "qclhip_grid_set_tiles": [_vp, _vp, ct.c_int, ct.c_int],
    "qclhip_batch_set_mode": [_vp, _vp, ct.c_int],
    "qclhip_batch_get_mode": [_vp],
    "qclhip_batch_autotune": [_vp, _vp, ct.POINTER(_vp), ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp),
                              ct.POINTER(_vp), _dp, ct.POINTER(_vp), ct.POINTER(_vp), ct.POINTER(_vp), _dp],
    "qclhip_batch_destroy": [_vp],
    "qclhip_batch_current_green": [_vp, _vp, ct.POINTER(_vp), ct.c_int, ct.POINTER(_vp), ct.POINTER(_vp),
                                   ct.POINTER(_vp), _dp, ct.POINTER(_vp), ct.POINTER(_vp), ct.POINTER(_vp), _dp],
}

_lib = None


def declared_symbols():
    return sorted(_SIGNATURES)


def load(path: str = LIB_PATH):
    """Load the library (without touching the GPU) and set the signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"libqclhip.so not found at {path}; build it with `python qclnegf.jl_amd/build.py` "
            "(the HIP backend has no CPU fallback)")
    lib = ct.CDLL(path)
    for name, argtypes in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = ct.c_char_p if name == "qclhip_ctx_last_error" else ct.c_int
    _lib = lib
    return lib


def check(rc: int, ctx=None, what: str = "") -> None:
    """Map a C-ABI return code to the reference's exception types."""
    if rc == OK:
        return
    msg = ""
    if ctx is not None and _lib is not None:
        raw = _lib.qclhip_ctx_last_error(ctx)
        msg = raw.decode() if raw else ""
    text = f"{what}: {msg}" if what else msg
    if rc > 0:
        raise SingularException(rc, f"({what})")
    if rc == ERR_ARGUMENT:
        raise ArgumentError(text)
    if rc == ERR_DIMENSION:
        raise DimensionMismatch(text)
    if rc == ERR_DOMAIN:
        raise DomainError(text)
    if rc == ERR_UNSUPPORTED:
        raise UnsupportedError(text)
    raise QCLHipError(f"runtime failure ({rc}) {text}")
