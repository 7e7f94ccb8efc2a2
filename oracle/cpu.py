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
