"""ctypes access to the C++/OpenMP CPU baseline (oracle/cpu_baseline.cpp).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Pinned against the
NumPy oracle (oracle/sv.py) in tests/test_oracle_cpu_baseline.py.
"""

import ctypes as C
import os

import numpy as np

_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "libcpu_baseline.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(_LIB)
        vp = C.c_void_p
        _lib.cpu_hea_grad.argtypes = [C.c_int, C.c_int, vp, vp, C.c_int, vp, vp]
        _lib.cpu_hea_grad.restype = C.c_int
        _lib.cpu_hea_grad_c128.argtypes = [C.c_int, C.c_int, vp, vp, C.c_int, vp, vp]
        _lib.cpu_hea_grad_c128.restype = C.c_int
        _lib.cpu_dm_forward.argtypes = [C.c_int, C.c_int, vp, C.c_double, C.c_double, C.c_int, vp]
        _lib.cpu_dm_forward.restype = C.c_int
        _lib.cpu_pulse_grad.argtypes = [C.c_int, vp, C.c_int, C.c_double, C.c_double, vp, C.c_int, vp, vp]
        _lib.cpu_pulse_grad.restype = C.c_int
    return _lib


def hea_grad(theta: np.ndarray, features=None, threads: int = 0, c128: bool = False):
    """(expvals (n,), grads (layers*n*3,)) of sum_q <Z_q> for the ansatz,
    computed in complex64 (or complex128 with ``c128``)."""
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    layers, n, _ = theta.shape
    f = None if features is None else np.ascontiguousarray(features, dtype=np.float64)
    ev = np.zeros(n)
    g = np.zeros(theta.size)
    fn = lib().cpu_hea_grad_c128 if c128 else lib().cpu_hea_grad
    fn(n, layers, theta.ctypes.data, None if f is None else f.ctypes.data, threads, ev.ctypes.data,
                       g.ctypes.data)
    return ev, g


def dm_forward(theta: np.ndarray, p_dep: float, gamma: float, threads: int = 0) -> float:
    """Tr(Z_0 rho) of the noisy density-matrix ansatz (complex64)."""
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    layers, n, _ = theta.shape
    out = np.zeros(1)
    lib().cpu_dm_forward(n, layers, theta.ctypes.data, p_dep, gamma, threads, out.ctypes.data)
    return float(out[0])


def pulse_grad(params: np.ndarray, omega: np.ndarray, n_steps: int, T: float, J: float, V: np.ndarray,
               threads: int = 0):
    """(infidelity (S,), gradient (S, 4)) of the two-qubit DRAG pulse sets,
    Magnus-2 with one expm per step, FP64."""
    P = np.ascontiguousarray(np.concatenate([params, omega], axis=1), dtype=np.float64)
    Vh = np.ascontiguousarray(np.stack([V.real, V.imag], axis=-1).reshape(-1))
    S = P.shape[0]
    inf = np.zeros(S)
    g = np.zeros((S, 4))
    lib().cpu_pulse_grad(S, P.ctypes.data, n_steps, T, J, Vh.ctypes.data, threads, inf.ctypes.data, g.ctypes.data)
    return inf, g
