"""Optimal-control cost functions (jaqsi_amd.qoc, mirrors jaqsi/qoc.py) on
the CPU: the Cost wrapper, the registry and the parameter-only costs, with
the reference's own cases (tests/test_qoc.py:32-262)."""

import math

import numpy as np
import pytest
import torch


@pytest.fixture(autouse=True)
def _restore_precision():
    import jaqsi_amd as J

    yield
    J.enable_x64(False)


def test_cost_wrapper():
    from jaqsi_amd.qoc import Cost

    assert float(Cost(cost=lambda p: p.sum(), weight=2.0)(torch.tensor([1.0, 3.0]))) == 8.0
    assert float(Cost(cost=lambda p: (p[0], p[1]), weight=(0.5, 0.25))(torch.tensor([4.0, 8.0]))) == 4.0
    assert float(Cost(cost=lambda p, offset=0.0: p.sum() + offset, ckwargs={"offset": 10.0})(torch.tensor([1.0, 2.0]))) == 13.0
    c = Cost(cost=lambda p: p[0]) + Cost(cost=lambda p: p[1])
    assert float(c(torch.tensor([3.0, 7.0]))) == 10.0
    assert float((Cost(cost=lambda p: p.sum(), weight=2.0) + None)(torch.tensor([1.0, 2.0]))) == 6.0
    with pytest.raises(ValueError):
        Cost(cost=lambda p: (p[0],), weight=(0.5, 0.5))(torch.tensor([1.0]))


def test_registry():
    from jaqsi_amd.qoc import CostFnRegistry, fidelity_cost_fn

    assert {"fidelity", "unitary", "pulse_width", "evolution_time", "spectral_density"} <= set(CostFnRegistry.available())
    assert CostFnRegistry.get("fidelity")["fn"] is fidelity_cost_fn
    assert CostFnRegistry.get("fidelity")["default_weight"] == (0.5, 0.5)
    with pytest.raises(ValueError, match="Unknown cost function"):
        CostFnRegistry.get("unknown")
    assert CostFnRegistry.parse_cost_arg("pulse_width:0.3") == ("pulse_width", 0.3)
    assert CostFnRegistry.parse_cost_arg("fidelity:0.6,0.2") == ("fidelity", (0.6, 0.2))
    assert CostFnRegistry.parse_cost_arg("evolution_time") == ("evolution_time", 1.0)
    for arg, match in (("bogus:0.5", "Unknown cost function"), ("fidelity:0.5", "expects 2 weight"),
                       ("pulse_width:0.3,0.2", "expects 1 weight")):
        with pytest.raises(ValueError, match=match):
            CostFnRegistry.parse_cost_arg(arg)


def test_parameter_costs():
    from jaqsi_amd.qoc import evolution_time_cost_fn, pulse_width_cost_fn, spectral_density_cost_fn

    assert float(pulse_width_cost_fn(torch.tensor([10.0, 2.5, 0.8]), envelope="gaussian")) == 2.5
    assert float(pulse_width_cost_fn(torch.tensor([0.5]), envelope="general")) == 0.0
    assert float(evolution_time_cost_fn(torch.tensor([5.0, 1.0, 1.0]), t_target=1.0)) == 0.0
    assert math.isclose(float(evolution_time_cost_fn(torch.tensor([5.0, 1.0, 1.5]), t_target=1.0)), 0.25)
    assert math.isclose(float(evolution_time_cost_fn(torch.tensor([0.0, 3.0]), 2.0)),
                        float(evolution_time_cost_fn(torch.tensor([0.0, 1.0]), 2.0)))
    p = torch.tensor([5.0, 1.0, 2.0], dtype=torch.float64, requires_grad=True)
    evolution_time_cost_fn(p, t_target=1.0).backward()
    assert p.grad[0] == 0 and p.grad[1] == 0 and p.grad[2] != 0
    sd = spectral_density_cost_fn
    assert float(sd(torch.tensor([0.5]), envelope="general")) == 0.0
    g = sd(torch.tensor([1.0, 0.3, 2.0], dtype=torch.float64), envelope="gaussian")
    assert g < sd(torch.tensor([1.0, 1.0, 2.0], dtype=torch.float64), envelope="square")
    assert 0.0 <= float(sd(torch.tensor([1.0, 0.5, 1.0], dtype=torch.float64), envelope="gaussian")) <= 1.0
    assert sd(torch.tensor([1.0, 0.5, 2.0], dtype=torch.float64), envelope="gaussian") < \
        sd(torch.tensor([1.0, 0.05, 2.0], dtype=torch.float64), envelope="gaussian")
    without = sd(torch.tensor([1.0, 0.0, 0.3, 2.0], dtype=torch.float64), envelope="drag")
    assert math.isclose(float(without), float(g), rel_tol=1e-12)
    assert sd(torch.tensor([1.0, 0.5, 0.3, 2.0], dtype=torch.float64), envelope="drag") > without
    q = torch.tensor([1.0, 0.3, 2.0], dtype=torch.float64, requires_grad=True)
    sd(q, envelope="gaussian").backward()
    assert torch.any(q.grad != 0)


def test_rotation_angle_sample():
    from jaqsi_amd.qoc import _sample_rotation_angles

    assert _sample_rotation_angles(1).tolist() == [0.0]
    w = _sample_rotation_angles(9).numpy()
    assert len(w) == 9 and w.min() >= 0 and w.max() < 2 * np.pi
    band = w[6:]
    np.testing.assert_allclose(band, np.pi / 2 + np.pi * np.arange(3) / 3)
    np.testing.assert_allclose(w[:6], 2 * np.pi * np.arange(6) / 6)
