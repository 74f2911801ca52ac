"""ctypes binding of libtcm.so (the C ABI declared in include/tcm.h).

The library is mandatory: importing this module raises if ``libtcm.so`` is
missing, and every compute entry point raises ``RuntimeError`` when the HIP
call fails (including "no HIP device").  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# TCM_LIBRARY selects another build of the same library (e.g. a debug build)
LIB_PATH = os.environ.get("TCM_LIBRARY") or os.path.join(HERE, "libtcm.so")

TCM_C64 = 0
TCM_C128 = 1

# tcm_gate_kind
G_MAT, G_RX, G_RY, G_RZ, G_RXX, G_RYY, G_RZZ, G_PHASE, G_EXP1, G_BARRIER, G_EXPG = range(11)

E_ARG, E_QUBIT, E_DUPQUBIT, E_WORKSPACE, E_UNSUPPORTED, E_NODEVICE, E_NONFINITE = (
    -1, -2, -3, -4, -5, -6, -7)


class TcmGate(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("nq", C.c_int32),
        ("q", C.c_int32 * 5),
        ("param", C.c_int32),
        ("theta", C.c_double),
        ("scale", C.c_double),
        ("mat", C.POINTER(C.c_double)),
    ]


class TcmOpts(C.Structure):
    _fields_ = [
        ("fuse", C.c_int32),
        ("tile_bits", C.c_int32),
        ("adj_tile_bits", C.c_int32),
        ("ckpt_every", C.c_int32),
        ("reg_slots", C.c_int32),
        ("mfma", C.c_int32),
        ("jit", C.c_int32),
        ("greedy", C.c_int32),
    ]


class TcmProgramInfo(C.Structure):
    _fields_ = [
        ("nqubits", C.c_int32),
        ("nparams", C.c_int32),
        ("ngates", C.c_int32),
        ("npasses", C.c_int32),
        ("nckpt", C.c_int32),
        ("tile_bits", C.c_int32),
        ("reserved", C.c_int32 * 2),
    ]


class TcmMultiDesc(C.Structure):
    _fields_ = [
        ("prog", C.c_void_p),
        ("params", C.c_void_p),
        ("param_stride", C.c_int64),
        ("state", C.c_void_p),
        ("batch", C.c_int64),
    ]


if not os.path.exists(LIB_PATH):
    raise ImportError(
        "libtcm.so not built (expected at %s); run `python tensorcircuit_amd/_build.py`" % LIB_PATH)

lib = C.CDLL(LIB_PATH)

_vp, _i32, _i64, _dp = C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_double)

# name -> (restype, argtypes); must match include/tcm.h exactly
SIGNATURES = {
    "tcm_abi_version": (C.c_int, []),
    "tcm_last_error": (C.c_char_p, []),
    "tcm_program_build": (C.c_int, [C.POINTER(TcmGate), _i32, _i32, _i32, C.POINTER(TcmOpts), C.POINTER(_vp)]),
    "tcm_program_free": (C.c_int, [_vp]),
    "tcm_program_get_info": (C.c_int, [_vp, C.POINTER(TcmProgramInfo)]),
    "tcm_program_describe": (C.c_int64, [_vp, _i32, _i32, C.c_char_p, _i64]),
    "tcm_program_compile": (C.c_int, [_vp, _i32, _i32, C.POINTER(C.c_double)]),
    "tcm_forward_workspace_bytes": (C.c_int, [_vp, _i32, _i64, C.POINTER(_i64)]),
    "tcm_forward": (C.c_int, [_vp, _i32, _vp, _i64, _vp, _i64, _i32, _vp, _vp, _i64, _vp]),
    "tcm_pauli_build": (C.c_int, [C.POINTER(C.c_int8), _dp, _i32, _i32, C.POINTER(_vp)]),
    "tcm_pauli_free": (C.c_int, [_vp]),
    "tcm_pauli_compile": (C.c_int, [_vp, _i32, C.POINTER(C.c_double)]),
    "tcm_pauli_workspace_bytes": (C.c_int, [_vp, _i32, _i64, C.POINTER(_i64)]),
    "tcm_expect_pauli_sum": (C.c_int, [_vp, _i32, _vp, _i64, _vp, _vp, _i64, _vp]),
    "tcm_expect_pauli_sum_noalloc": (C.c_int, [_vp, _i32, _vp, _i64, _vp, _vp]),
    "tcm_apply_pauli_sum": (C.c_int, [_vp, _i32, _vp, _vp, _i64, _vp, _vp]),
    "tcm_pauli_terms_workspace_bytes": (C.c_int, [_vp, _i64, C.POINTER(_i64)]),
    "tcm_expect_pauli_terms": (C.c_int, [_vp, _i32, _vp, _i64, _vp, _vp, _i64, _vp]),
    "tcm_apply_pauli_terms": (C.c_int, [_vp, _i32, _vp, _vp, _i64, _vp, _vp]),
    "tcm_csr_matvec": (C.c_int, [_i32, _i32, _vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    "tcm_inner_workspace_bytes": (C.c_int, [_i32, _i64, C.POINTER(_i64)]),
    "tcm_inner": (C.c_int, [_i32, _i32, _vp, _vp, _i64, _vp, _vp, _i64, _vp]),
    "tcm_probability": (C.c_int, [_i32, _i32, _vp, _i64, _vp, _vp]),
    "tcm_pauli_nnz": (C.c_int, [_vp, C.POINTER(_i64)]),
    "tcm_pauli_coo": (C.c_int, [_vp, _i32, _vp, _vp, _vp, _vp]),
    "tcm_csr_expect_workspace_bytes": (C.c_int, [_i32, C.POINTER(_i64)]),
    "tcm_csr_expect": (C.c_int, [_i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    "tcm_sample_workspace_bytes": (C.c_int, [_i32, _i64, C.POINTER(_i64)]),
    "tcm_sample": (C.c_int, [_i32, _i32, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    "tcm_sample_prob": (C.c_int, [_i32, _i32, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    "tcm_measure_workspace_bytes": (C.c_int, [_i32, _i64, C.POINTER(_i64)]),
    "tcm_measure": (C.c_int, [_i32, _i32, _vp, _i64, C.POINTER(_i32), _i32, _vp, _vp, _vp, _vp, _i64, _vp]),
    "tcm_adjoint_workspace_bytes": (C.c_int, [_vp, _vp, _i32, _i64, _i32, C.POINTER(_i64)]),
    "tcm_adjoint_energy_grad": (C.c_int, [_vp, _i32, _vp, _i64, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _vp]),
    "tcm_adjoint_vjp": (C.c_int, [_vp, _i32, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _vp]),
    "tcm_multi_workspace_bytes": (C.c_int, [C.POINTER(TcmMultiDesc), _i32, _i32, C.POINTER(_i64)]),
    "tcm_forward_multi": (C.c_int, [C.POINTER(TcmMultiDesc), _i32, _i32, _i32, _vp, _i64, _vp]),
    "tcm_multi_plan_build": (C.c_int, [C.POINTER(_vp), C.POINTER(_i64), _i32, _i32, C.POINTER(_vp),
                                       C.POINTER(C.c_double)]),
    "tcm_multi_plan_free": (C.c_int, [_vp]),
    "tcm_forward_multi_plan": (C.c_int, [_vp, C.POINTER(TcmMultiDesc), _i32, _i32, _vp, _i64, _vp]),
}

for _name, (_res, _args) in SIGNATURES.items():
    _f = getattr(lib, _name)  # AttributeError here = a symbol the header declares is missing
    _f.restype = _res
    _f.argtypes = _args

if lib.tcm_abi_version() != 1:
    raise ImportError("libtcm ABI version mismatch")


def last_error() -> str:
    msg = lib.tcm_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = "") -> None:
    """Map a tcm status code to the reference's exception types: argument
    errors -> ValueError (basecircuit.py:306-307 style), HIP / device
    errors -> RuntimeError."""
    if rc == 0:
        return
    msg = "%s: %s (code %d)" % (what, last_error(), rc)
    if rc in (E_ARG, E_QUBIT, E_DUPQUBIT, E_NONFINITE):
        raise ValueError(msg)
    if rc == E_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)
