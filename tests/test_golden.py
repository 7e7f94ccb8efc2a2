"""The committed golden fixtures are reproduced by the oracle (CPU)."""

import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return dict(np.load(os.path.join(HERE, name + ".npz")))


@pytest.mark.parametrize("name", ["c1_hea", "dm_small"])
def test_oracle_reproduces_fixture(name):
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    fresh = mg.MAKERS[name]()
    ref = _load(name)
    assert set(fresh) == set(ref)
    for k in ref:
        np.testing.assert_allclose(fresh[k], ref[k], rtol=1e-12, atol=1e-14, err_msg=k)


def test_cpu_baseline_port_on_c1_fixture():
    from oracle import cpu

    g = _load("c1_hea")
    ev, grads = cpu.hea_grad(g["theta"], None, threads=2)
    np.testing.assert_allclose(ev, g["expvals"], atol=2e-6)
    np.testing.assert_allclose(grads, g["grads"], atol=2e-5)


def test_c5_fixture_consistency():
    from oracle import pulse

    g = _load("c5_small")
    V = pulse.target_rx_half_pi()
    U = pulse.propagate(g["params"][0], g["omega"][0], pulse.J_C5, pulse.T_C5, int(g["n_steps"]))
    np.testing.assert_allclose(U, g["U"][0], atol=1e-13)
    assert abs(pulse.infidelity(U, V) - g["infidelity"][0]) < 1e-13
