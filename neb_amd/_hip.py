"""ctypes binding of libnebhip.so (the C ABI declared in include/neb_hip.h).

The library is mandatory: if it is missing every entry point raises, there is
no CPU or PyTorch fallback.  Error codes map to the exception classes the
reference backend raises (ValueError for shapes/arguments, FloatingPointError
for nonfinite values, RuntimeError for driver errors; HB2P:3299-3305,
VCORE:60-61 in experiments/metaball/_backend/flow/cases/metaball/velocity/
core.py).
"""
from __future__ import annotations

import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libnebhip.so")
ABI_VERSION = 16

F32, BF16, F64 = 0, 1, 2
TANH, SIN = 0, 1
LIFT_2D_VORTICITY, LIFT_3D_CURL, LIFT_2D_HESSIAN, LIFT_2D_STREAM, LIFT_2D_STREAM_JET1 = 2, 3, 4, 5, 6
DIM_2D_JET1 = 12
CROSS_F32, CROSS_BF16, CROSS_BF16X3, CROSS_F64 = 0, 1, 2, 3

_i64, _dbl, _int, _vp, _sz = C.c_int64, C.c_double, C.c_int, C.c_void_p, C.c_size_t

# name -> (restype, argtypes)
SIGNATURES = {
    "neb_abi_version": (_int, []),
    "neb_last_error": (C.c_char_p, []),
    "neb_feature_jets": (_int, [_int, _int, _i64, _vp, _int, _vp, _vp, _vp, _i64, _vp]),
    "neb_jet_layer_forward": (_int, [_int, _int, _int, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _dbl,
                                     _vp, _vp, _i64, _vp]),
    "neb_jet_output_jets": (_int, [_int, _int, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    "neb_jet_output_jets_enveloped": (_int, [_int, _int, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _vp,
                                             _i64, _vp]),
    "neb_envelope_reverse": (_int, [_int, _int, _i64, _i64, _vp, _vp, _i64, _vp, _i64, _vp]),
    "neb_jet_output_fields": (_int, [_int, _int, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _int,
                                     _vp]),
    "neb_jet_output_fields_split": (_int, [_int, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _vp, _vp, _i64,
                                           _int, _vp]),
    "neb_split_bf16": (_int, [_int, _vp, _i64, _i64, _i64, _vp, _vp, _i64, _vp]),
    "neb_center_columns": (_int, [_int, _vp, _vp, _i64, _i64, _i64, _vp, _vp]),
    "neb_join_bf16": (_int, [_int, _vp, _vp, _i64, _i64, _i64, _vp, _i64, _vp]),
    "neb_jet_layer_forward_x3": (_int, [_int, _int, _i64, _i64, _i64, _vp, _vp, _i64, _vp, _vp, _i64, _vp, _dbl,
                                        _vp, _vp, _vp, _vp, _i64, _vp]),
    "neb_jet_output_fields_x3": (_int, [_int, _i64, _i64, _i64, _vp, _vp, _i64, _vp, _vp, _i64, _vp, _vp, _vp, _vp,
                                        _i64, _int, _vp]),
    "neb_jet_output_backward": (_int, [_int, _int, _int, _i64, _i64, _i64, _vp, _vp, _i64, _int, _int, _vp, _vp,
                                       _i64, _vp, _i64, _dbl, _vp, _i64, _vp]),
    "neb_jet_output_backward_blocks": (_int, [_int, _int, _int, _i64, _i64, _i64, _vp, _vp, _i64, _int, _int, _vp,
                                              _vp, _i64, _vp, _i64, _dbl, _vp, _i64, _vp]),
    "neb_jet_layer_backward": (_int, [_int, _int, _int, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _dbl,
                                      _vp, _i64, _vp]),
    "neb_jet_layer_forward_blocks": (_int, [_int, _int, _int, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _dbl,
                                            _vp, _vp, _i64, _vp]),
    "neb_jet_layer_forward_x3_blocks": (_int, [_int, _int, _i64, _i64, _i64, _vp, _vp, _i64, _vp, _vp, _i64, _vp,
                                               _dbl, _vp, _vp, _vp, _vp, _i64, _vp]),
    "neb_jet_layer_backward_blocks": (_int, [_int, _int, _int, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64,
                                             _dbl, _vp, _i64, _vp]),
    "neb_cross_blocks_workspace_size": (_sz, [_i64, _i64, _i64]),
    "neb_cross_blocks": (_int, [_int, _int, _vp, _vp, _vp, _vp, _vp, _vp, _int, _int, _i64, _i64, _dbl, _dbl, _vp,
                                _i64, _vp, _sz, _vp]),
    "neb_mode_output_fields_x3": (_int, [_int, _i64, _i64, _i64, _vp, _vp, _i64, _vp, _vp, _i64, _vp, _vp, _vp,
                                         _vp, _i64, _vp]),
    "neb_stack_channels": (_int, [_vp, _vp, _i64, _i64, _i64, _int, _vp, _vp, _vp, _i64, _vp]),
    "neb_gemm_split": (_int, [_i64, _i64, _i64, _vp, _vp, _i64, _vp, _vp, _i64, _vp, _i64, _vp]),
    "neb_jet_act_reverse": (_int, [_int, _int, _int, _i64, _i64, _vp, _i64, _int, _vp, _vp, _i64, _dbl, _vp, _i64,
                                   _vp]),
    "neb_cross_workspace_size": (_sz, [_i64, _i64, _i64, _int]),
    "neb_cross": (_int, [_int, _int, _vp, _vp, _vp, _vp, _vp, _vp, _int, _int, _i64, _i64, _vp, _int, _dbl,
                         _dbl, _vp, _i64, _vp, _sz, _vp]),
    "neb_gram": (_int, [_int, _vp, _i64, _i64, _i64, _int, _dbl, _vp, _dbl, _vp, _i64, _vp, _sz, _vp]),
    "neb_gram_split": (_int, [_vp, _vp, _i64, _i64, _i64, _dbl, _dbl, _vp, _i64, _vp, _sz, _vp]),
    "neb_gram_split_noalloc": (_int, [_vp, _vp, _i64, _i64, _i64, _dbl, _dbl, _vp, _i64, _vp]),
    "neb_set_cross_occupancy": (_int, [_int]),
    "neb_cross_modes_workspace_size": (_sz, [_i64, _i64, _i64]),
    "neb_cross_modes": (_int, [_int, _vp, _vp, _vp, _vp, _vp, _vp, _int, _int, _i64, _i64, _dbl, _dbl, _vp, _vp, _sz,
                               _vp]),
    "neb_colsum_workspace_size": (_sz, [_i64, _i64]),
    "neb_colsum": (_int, [_int, _vp, _i64, _i64, _i64, _dbl, _vp, _vp, _sz, _vp]),
    "neb_sumsq": (_int, [_int, _vp, _i64, _vp, _vp, _sz, _vp]),
    "neb_adam": (_int, [_int, _vp, _vp, _vp, _vp, _i64, _dbl, _dbl, _dbl, _dbl, _dbl, _i64, _vp]),
    "neb_adam_gated": (_int, [_int, _vp, _vp, _vp, _vp, _i64, _dbl, _dbl, _dbl, _dbl, _dbl, _i64, _vp, _vp, _vp]),
    "neb_galerkin_workspace_size": (_sz, [_i64, _i64]),
    "neb_galerkin_rhs2d": (_int, [_int, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _dbl, _vp, _vp, _sz, _vp]),
    "neb_stage_mode_block": (_int, [_int, _vp, _i64, _i64, _i64, _i64, _i64, _int, _int, _vp, _vp, _i64, _vp, _i64,
                                    _vp]),
    "neb_stage_net": (_int, [_int, _vp, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "neb_stage": (_int, [_int, _int, _vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64, _int, _dbl, _vp]),
    "neb_sym_shift": (_int, [_i64, _vp, _i64, _dbl, _vp, _i64, _vp]),
    "neb_spectral_workspace_size": (_sz, [_i64]),
    "neb_spectral_terms": (_int, [_i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _dbl, _dbl, _dbl, _vp, _i64, _vp, _vp,
                                  _vp, _sz, _vp]),
    "neb_spectral_adjoint": (_int, [_i64, _vp, _i64, _vp, _i64, _dbl, _vp, _i64, _vp]),
    "neb_spd_inverse_workspace_size": (_sz, [_i64]),
    "neb_spd_inverse": (_int, [_i64, _vp, _i64, _vp, _i64, _vp, _vp, _sz, _vp]),
    "neb_spd_inverse_chol": (_int, [_i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _sz, _vp]),
}

_lib = None


def load():
    """Load libnebhip.so (once) and bind every exported symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m neb_amd.build` "
                           "(there is no fallback path)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.neb_abi_version() != ABI_VERSION:
        raise RuntimeError(f"libnebhip ABI {lib.neb_abi_version()} != {ABI_VERSION}")
    _lib = lib
    return lib


_EXC = {1: ValueError, 2: FloatingPointError, 3: RuntimeError, 4: ValueError}


def check(rc: int) -> None:
    if rc:
        msg = _lib.neb_last_error().decode(errors="replace")
        raise _EXC.get(rc, RuntimeError)(msg)


_recorder = None


def set_launch_recorder(recorder) -> None:
    """Install (or, with None, remove) a per-launch hook: recorder(name, args)
    is called before each entry point and must return a callable that is
    called after it (a timing harness records stream events there).  The
    default is no hook."""
    global _recorder
    _recorder = recorder


def call(name: str, *args) -> None:
    if _recorder is None:
        check(getattr(load(), name)(*args))
        return
    done = _recorder(name, args)
    check(getattr(load(), name)(*args))
    done()


def ptr_array(ptrs):
    return (C.c_void_p * len(ptrs))(*ptrs)


def i64_array(vals):
    return (C.c_int64 * len(vals))(*vals)


def int_array(vals):
    return (C.c_int * len(vals))(*vals)


def dbl_array(vals):
    return (C.c_double * len(vals))(*vals)
