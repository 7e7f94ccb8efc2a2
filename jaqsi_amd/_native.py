"""ctypes binding of libjaqsi_amd.so (the C-ABI in include/jaqsi_amd.h).

The library is loaded from the package directory.  There is no fallback:
if it is missing, every simulation entry point raises, so a result never
silently comes from anything but the HIP kernels.
"""

from __future__ import annotations

import ctypes as C
import os
import threading
from typing import Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjaqsi_amd.so")

JQ_OK = 0
JQ_ERR_ARG = 1
JQ_ERR_UNSUPPORTED = 2
JQ_ERR_ALLOC = 3
JQ_ERR_DEVICE = 4
JQ_ERR_NONFINITE = 5

JQ_M_ROT, JQ_M_PHASE, JQ_M_MATRIX = 0, 1, 2
JQ_MF_CONJ, JQ_MF_SUPEROP, JQ_MF_GRAD, JQ_MF_UNITARY = 1, 2, 4, 8
JQ_S_DIAGONAL, JQ_S_PERMUTATION, JQ_S_NONUNITARY = 1, 2, 4
JQ_OBS_NONE, JQ_OBS_SV_DIAG, JQ_OBS_DM_DIAG, JQ_OBS_SV_PARITY = 0, 1, 2, 3
MAX_WIRES = 6
JQ_COND_MAX = 4.0
JQ_TUNE_NO_PERSIST_BWD = 1
JQ_TUNE_NO_DMA_FWD = 2
JQ_TUNE_NARROW_BWD = 4
JQ_TUNE_BWD16 = 8
JQ_NOALLOC_SCRATCH_BYTES = 32 << 20


class Member(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("flags", C.c_int32),
        ("nctrl", C.c_int32),
        ("word", C.c_int32),
        ("param", C.c_int32),
        ("matrix", C.c_int32),
        ("value", C.c_double),
    ]


class Op(C.Structure):
    _fields_ = [
        ("nwires", C.c_int32),
        ("wires", C.c_int32 * MAX_WIRES),
        ("structure", C.c_int32),
        ("controls", C.c_int32),
        ("perm", C.c_int8 * 64),
        ("member_begin", C.c_int32),
        ("member_end", C.c_int32),
        ("cond", C.c_double),
    ]


class PlanDesc(C.Structure):
    _fields_ = [
        ("n_wires", C.c_int32),
        ("n_ops", C.c_int32),
        ("ops", C.POINTER(Op)),
        ("n_members", C.c_int32),
        ("members", C.POINTER(Member)),
        ("n_params", C.c_int32),
        ("n_matrix_elems", C.c_int32),
        ("dtype", C.c_int32),
        ("tile_bits", C.c_int32),
        ("checkpoints", C.c_int32),
        ("low_bits", C.c_int32),
        ("variant_flags", C.c_int32),
        ("fold_tail", C.c_int32),
    ]


class Group(C.Structure):
    """jq_group: one structure group of a ragged batch."""

    _fields_ = [
        ("plan", C.c_void_p),
        ("batch", C.c_int32),
        ("psi0_batched", C.c_int32),
        ("angles", C.c_void_p),
        ("matrices", C.c_void_p),
        ("psi0", C.c_void_p),
        ("ct", C.c_void_p),
        ("expvals", C.c_void_p),
        ("grad_angles", C.c_void_p),
        ("state", C.c_void_p),
    ]


class PulseDesc(C.Structure):
    _fields_ = [
        ("n_sets", C.c_int32),
        ("n_steps", C.c_int32),
        ("t_total", C.c_double),
        ("coupling_j", C.c_double),
        ("envelope", C.c_int32),
        ("reserved", C.c_int32),
    ]


JQ_ENV_DRAG_GAUSSIAN = 1


class EvolveDesc(C.Structure):
    _fields_ = [
        ("n_gates", C.c_int32),
        ("dim", C.c_int32),
        ("n_terms", C.c_int32),
        ("n_factors", C.c_int32),
        ("hmats_batched", C.c_int32),
        ("reserved", C.c_int32),
    ]

# plan-specialised kernels (jq_plan_jit)
JQ_JIT_OFF, JQ_JIT_ASYNC, JQ_JIT_SYNC, JQ_JIT_QUERY = 0, 1, 2, -1
JQ_JIT_NONE, JQ_JIT_COMPILING, JQ_JIT_COMPILED, JQ_JIT_LOADED, JQ_JIT_FAILED = 0, 1, 2, 3, -1
JQ_JIT_MIN_WIRES = 20


_lib = None
_lock = threading.Lock()

EXPORTS = [
    "jq_plan_create",
    "jq_plan_destroy",
    "jq_plan_info",
    "jq_plan_program",
    "jq_plan_tables",
    "jq_plan_jit",
    "jq_sv_forward",
    "jq_sv_adjoint",
    "jq_sv_forward_noalloc",
    "jq_sv_forward_ckpt",
    "jq_sv_backward",
    "jq_sv_adjoint_groups",
    "jq_sv_probs",
    "jq_sv_outer",
    "jq_dm_diag",
    "jq_expval_matrix",
    "jq_pulse_propagate",
    "jq_evolve",
    "jq_evolve_vjp",
    "jq_timing_enable",
    "jq_timing_read",
    "jq_last_error",
    "jq_version",
]


class NativeError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m jaqsi_amd.build` "
                "(the HIP extension is required; there is no CPU fallback)"
            )
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.jq_plan_create.argtypes = [C.POINTER(PlanDesc), C.POINTER(vp)]
        L.jq_plan_destroy.argtypes = [vp]
        L.jq_plan_info.argtypes = [vp] + [C.POINTER(C.c_int32)] * 5
        L.jq_plan_program.argtypes = [vp, C.POINTER(C.c_int32), C.c_int64, C.POINTER(C.c_int64)]
        L.jq_plan_tables.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32), C.c_int64, C.POINTER(C.c_int64)]
        L.jq_plan_jit.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32)]
        L.jq_sv_forward.argtypes = [vp, C.c_int32, vp, vp, vp, C.c_int32, vp, C.c_int32, C.c_int32, vp, vp, vp, vp]
        L.jq_sv_adjoint.argtypes = [vp, C.c_int32, vp, vp, vp, C.c_int32, vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp, vp, vp, vp]
        L.jq_sv_forward_noalloc.argtypes = [vp, C.c_int32, vp, vp, vp, C.c_int32, vp, C.c_int32, C.c_int32, vp, vp,
                                            vp, vp]
        L.jq_sv_forward_ckpt.argtypes = [vp, C.c_int32, vp, vp, vp, C.c_int32, vp, C.c_int32, C.c_int32, vp, vp, vp,
                                         C.POINTER(C.c_int64), vp]
        L.jq_sv_backward.argtypes = [vp, C.c_int64, C.c_int32, vp, vp, vp, C.c_int32, vp, C.c_int32, C.c_int32, vp, vp,
                                     vp, vp, vp, vp, vp, vp]
        L.jq_sv_adjoint_groups.argtypes = [C.c_int32, C.POINTER(Group), C.c_int32, C.c_int32, vp, vp, vp]
        L.jq_pulse_propagate.argtypes = [C.POINTER(PulseDesc), vp, vp, vp, vp, vp, vp]
        L.jq_evolve.argtypes = [C.POINTER(EvolveDesc), vp, vp, vp, vp]
        L.jq_evolve_vjp.argtypes = [C.POINTER(EvolveDesc), vp, vp, vp, vp, vp]
        L.jq_sv_probs.argtypes = [C.c_int32, C.c_int32, C.c_int32, vp, vp, vp]
        L.jq_sv_outer.argtypes = [C.c_int32, C.c_int32, C.c_int32, vp, vp, vp]
        L.jq_dm_diag.argtypes = [C.c_int32, C.c_int32, C.c_int32, vp, vp, vp]
        L.jq_expval_matrix.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp, C.c_int32, vp, vp, vp, vp]
        L.jq_timing_enable.argtypes = [C.c_int32]
        L.jq_timing_read.argtypes = [C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int32)]
        L.jq_last_error.restype = C.c_char_p
        L.jq_version.restype = C.c_char_p
        for name in EXPORTS:
            if name not in ("jq_last_error", "jq_version"):
                getattr(L, name).restype = C.c_int
        _lib = L
        return L


def check(rc: int) -> None:
    if rc != JQ_OK:
        msg = lib().jq_last_error().decode()
        if rc == JQ_ERR_ARG:
            raise ValueError(msg)
        if rc == JQ_ERR_UNSUPPORTED:
            raise NotImplementedError(msg)
        raise NativeError(rc, msg)


def ptr(t) -> Optional[int]:
    """Device (or host) address of a tensor / array, None for None."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    return t.data_ptr()


class Plan:
    """Owner of a native jq_plan (one fused, scheduled circuit structure)."""

    def __init__(self, n_wires, ops, members, n_params, n_matrix_elems, dtype, tile_bits=0, checkpoints=1,
                 low_bits=0, variant_flags=0, fold_tail=0):
        L = lib()
        self._ops = (Op * max(1, len(ops)))(*ops)
        self._members = (Member * max(1, len(members)))(*members)
        d = PlanDesc(n_wires, len(ops), self._ops, len(members), self._members, n_params, n_matrix_elems, dtype,
                     tile_bits, checkpoints, low_bits, variant_flags, int(fold_tail))
        self.low_bits = low_bits
        self.variant_flags = variant_flags
        self.fold_tail = int(fold_tail)
        h = C.c_void_p()
        check(L.jq_plan_create(C.byref(d), C.byref(h)))
        self.handle = h
        self.n_wires = n_wires
        self.n_params = n_params
        self.n_matrix_elems = n_matrix_elems
        self.dtype = dtype

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib is not None:
            _lib.jq_plan_destroy(h)
            self.handle = None

    def info(self) -> dict:
        vals = [C.c_int32() for _ in range(5)]
        check(lib().jq_plan_info(self.handle, *[C.byref(v) for v in vals]))
        keys = ["n_passes", "n_substeps", "n_blocks", "n_grad_reals", "tile_bits"]
        return {k: v.value for k, v in zip(keys, vals)}

    def jit(self, mode: int = JQ_JIT_QUERY) -> int:
        """Select the plan-specialised kernels (JQ_JIT_OFF / _ASYNC / _SYNC);
        returns the compile phase (JQ_JIT_NONE ... JQ_JIT_LOADED, _FAILED)."""
        ph = C.c_int32()
        check(lib().jq_plan_jit(self.handle, mode, C.byref(ph)))
        return ph.value

    def program(self) -> np.ndarray:
        n = C.c_int64()
        check(lib().jq_plan_program(self.handle, None, 0, C.byref(n)))
        out = np.zeros(max(1, n.value), dtype=np.int32)
        check(lib().jq_plan_program(self.handle, out.ctypes.data_as(C.POINTER(C.c_int32)), out.size, C.byref(n)))
        return out[: n.value]

    def fold(self) -> Tuple[int, np.ndarray, int]:
        """(ops folded, rows, flip) of the tail measured through the
        observables: the final bit of wire w is the parity of the wires in
        rows[w] (bit v = wire v) on the vector the passes compute, XOR bit w
        of flip."""
        t = self.table(6).astype(np.int64) & 0xFFFFFFFF
        rows = t[3::2] | (t[4::2] << 32)
        return int(t[0]), rows.astype(np.uint64), int(t[1] | (t[2] << 32))

    def table(self, which: int) -> np.ndarray:
        n = C.c_int64()
        check(lib().jq_plan_tables(self.handle, which, None, 0, C.byref(n)))
        out = np.zeros(max(1, n.value), dtype=np.int32)
        check(lib().jq_plan_tables(self.handle, which, out.ctypes.data_as(C.POINTER(C.c_int32)), out.size, C.byref(n)))
        return out[: n.value]
