"""The parity bar of the GPU tests.

Every value and gradient the GPU returns is held to a relative error of
1e-5 in complex64 and 1e-12 in complex128 against the complex128 oracle.
Entries much smaller than the array's largest magnitude (an expectation
value near zero, a vanishing gradient component) cannot carry a relative
error of their own; for them the bar is the same fraction of the array's
largest magnitude.
"""

import numpy as np

RTOL = {False: 1e-5, True: 1e-12}


def assert_parity(got, want, x64, scale=None, err_msg=""):
    """|got - want| <= rtol * |want| + rtol * max|want| elementwise.

    ``scale`` replaces max|want| when the reference array itself is a slice
    of a larger result (e.g. one circuit of a batch)."""
    got = np.asarray(got)
    want = np.asarray(want)
    rtol = RTOL[bool(x64)]
    if scale is None:
        scale = float(np.abs(want).max()) if want.size else 0.0
    np.testing.assert_allclose(got, want, rtol=rtol, atol=rtol * scale, err_msg=err_msg)


def max_rel_err(got, want):
    """Largest deviation relative to the reference's largest magnitude."""
    got = np.asarray(got)
    want = np.asarray(want)
    s = float(np.abs(want).max()) if want.size else 0.0
    return float(np.abs(got - want).max() / s) if s > 0 else float(np.abs(got - want).max())
