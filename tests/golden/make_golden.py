"""Write the golden fixtures of this directory from the complex128 / float64
oracle (the reference itself cannot be imported in this environment: its
jax 0.11 / diffrax / optax pins are absent and it needs Python >= 3.12).

    python tests/golden/make_golden.py

Each fixture holds the inputs (seeded as the benchmark workloads) and the
expected outputs; tests/test_golden.py checks that the oracle still
reproduces them and tests/test_gpu_golden.py checks the GPU path against them.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import dm, pulse, sv  # noqa: E402


def c1():
    """10 qubits, 5 layers, 150 angles U[0, 2pi) seed 1001; Z expvals and
    the adjoint gradient of their sum."""
    n, layers = 10, 5
    theta = np.random.default_rng(1001).uniform(0, 2 * np.pi, (layers, n, 3))
    ops = sv.hea_ops(theta, n)
    vals, grads, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
    return dict(theta=theta, expvals=vals, grads=sv.flat_grads(ops, grads))


def c3_head(count=6):
    """The first circuits of the 4096 x 14-qubit ragged batch (seed 1003)."""
    B, n = 4096, 14
    rng = np.random.default_rng(1003)
    depths = rng.integers(2, 13, size=B)
    feats = rng.uniform(-np.pi, np.pi, (B, n))
    thetas = [rng.uniform(0, 2 * np.pi, (int(d), n, 3)) for d in depths]
    out = {"depths": depths[:count], "feats": feats[:count]}
    for b in range(count):
        ops = sv.hea_ops(thetas[b], n, features=feats[b])
        vals, grads, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
        out[f"theta_{b}"] = thetas[b]
        out[f"expvals_{b}"] = vals
        out[f"grads_{b}"] = sv.flat_grads(ops, grads)
    return out


def dm_small():
    """Noisy ansatz on 3 qubits, 2 layers: DepolarizingChannel(0.01) and
    AmplitudeDamping(0.005) after every gate on each touched wire; Tr(Z0 rho)
    and its exact parameter-shift gradient."""
    n = 3
    theta = np.random.default_rng(1004).uniform(0, 2 * np.pi, (2, n, 3))
    v, g = dm.value_and_shift_grad(dm.noisy_hea(theta, n, 0.01, 0.005), n, [(sv.PZ, (0,))], np.array([1.0]))
    return dict(theta=theta, value=np.array(v), grads=g)


def c5_small(count=4, n_steps=200):
    """The first pulse parameter sets of seed 1005 at 200 Magnus-2 steps:
    propagators, infidelities and central-difference gradients."""
    params, omega = pulse.workload(count, 1005)
    V = pulse.target_rx_half_pi()
    U = np.stack([pulse.propagate(params[s], omega[s], pulse.J_C5, pulse.T_C5, n_steps) for s in range(count)])
    inf = np.array([pulse.infidelity(U[s], V) for s in range(count)])
    g = np.zeros((count, 4))
    for s in range(count):
        for k in range(4):
            h = 1e-5 * max(1.0, abs(params[s, k]))
            pp, pm = params[s].copy(), params[s].copy()
            pp[k] += h
            pm[k] -= h
            fp = pulse.infidelity(pulse.propagate(pp, omega[s], pulse.J_C5, pulse.T_C5, n_steps), V)
            fm = pulse.infidelity(pulse.propagate(pm, omega[s], pulse.J_C5, pulse.T_C5, n_steps), V)
            g[s, k] = (fp - fm) / (2 * h)
    return dict(params=params, omega=omega, n_steps=np.array(n_steps), U=U, infidelity=inf, grads=g)


MAKERS = {"c1_hea": c1, "c3_head": c3_head, "dm_small": dm_small, "c5_small": c5_small}


def main():
    for name, fn in MAKERS.items():
        path = os.path.join(HERE, name + ".npz")
        np.savez(path, **fn())
        print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
