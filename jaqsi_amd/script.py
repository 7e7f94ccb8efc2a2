"""``Script``: the user-facing circuit runner (mirrors jaqsi/script.py).

``Script(f, n_qubits).execute(type, obs, args=..., kwargs=..., in_axes=...,
shots=..., key=..., initial_state=...)`` records ``f`` on a tape and hands
it to :func:`jaqsi_amd.simulation.simulate_and_measure`.

Batched execution follows ``jax.vmap`` semantics as the reference does
(script.py:43-127, 527-744): ``in_axes`` names the batch axis of each
argument or ``None`` to broadcast.  The circuit is recorded once under
``torch.func.vmap``; the recording yields the shared circuit structure and
the batched gate parameters, which the GPU then runs as one launch.
"""

from __future__ import annotations

from typing import Callable, Hashable, List, Optional, Tuple

import numpy as np
import torch

from jaqsi_amd import lowering, simulation
from jaqsi_amd.gateset import Barrier
from jaqsi_amd.operations import Operation, _as_tensor
from jaqsi_amd.tape import RecordPerSample, recording, vectorized_recording


def _to_tensor(a):
    if isinstance(a, torch.Tensor):
        return a
    if isinstance(a, (np.ndarray, list, tuple, float, int)):
        return torch.as_tensor(np.asarray(a))
    return a


class Script:
    def __init__(self, f: Callable[..., None], n_qubits: Optional[int] = None) -> None:
        self.f = f
        self._n_qubits = n_qubits

    def __getstate__(self) -> dict:
        return dict(self.__dict__)

    def record(self, *args, **kwargs) -> List[Operation]:
        """Run the circuit function and return its tape (script.py:186-204)."""
        with recording() as tape:
            self.f(*args, **kwargs)
        return tape

    def execute(
        self,
        type: str = "expval",
        obs: Optional[List[Operation]] = None,
        *,
        args: tuple = (),
        kwargs: Optional[dict] = None,
        in_axes: Optional[Tuple] = None,
        shots: Optional[int] = None,
        key=None,
        initial_state=None,
        fingerprint: Optional[Hashable] = None,
    ):
        """Execute the circuit and return measurement results (script.py:226-335)."""
        if obs is None:
            obs = []
        if kwargs is None:
            kwargs = {}
        if initial_state is not None:
            nd = np.ndim(initial_state)
            if in_axes is None and nd != 1:
                raise ValueError(f"initial_state must be a 1D statevector when in_axes is None; got ndim={nd}.")
            if in_axes is not None and nd not in (1, 2):
                raise ValueError(
                    f"Batched initial_state must be 1D (broadcast) or 2D (one state per sample); got ndim={nd}."
                )
        if in_axes is not None:
            return self._execute_batched(type, obs, args, kwargs, in_axes, shots, key, initial_state)
        tape = self.record(*args, **kwargs)
        n_qubits = self._n_qubits or simulation.infer_n_qubits(tape, obs)
        use_density = simulation.has_noise(tape)
        return simulation.simulate_and_measure(
            tape, n_qubits, type, obs, use_density, shots=shots, key=key, initial_state=initial_state
        )

    def _execute_per_sample(self, type, obs, args, kwargs, in_axes, B, shots, key, initial_state):
        """Batched execution of circuits with pulse gates: the circuit is
        recorded once per sample (the propagators cannot be solved inside
        vmap), every pending evolution of all samples is solved together
        (one propagator launch per solver group, identical gates once), and
        the samples run as one batched simulation.  The samples must share
        one gate structure."""
        from jaqsi_amd.evolution import resolve_pending

        tapes = []
        for b in range(B):
            ab = [a.select(ax, b) if ax is not None else a for a, ax in zip(args, in_axes)]
            with recording() as tape:
                self.f(*ab, **kwargs)
            tapes.append([op for op in tape if not isinstance(op, Barrier)])
        resolve_pending([op for t in tapes for op in t])
        n_qubits = self._n_qubits or simulation.infer_n_qubits(tapes[0], obs)
        dens = simulation.has_noise(tapes[0])
        lower = lowering.lower_mixed if dens else lowering.lower_pure
        lows = [lower(t, n_qubits) for t in tapes]
        if any(low.key != lows[0].key for low in lows[1:]):
            raise ValueError("the samples of a batched execution must share one gate structure")
        dev = simulation.device()
        angles = torch.stack([lowering.angle_tensor(low.angle_values, None).to(dev) for low in lows])
        mats = None
        if lows[0].n_matrix_elems:
            mats = torch.stack([lowering.matrix_tensor(low.matrix_values, low.n_matrix_elems, None).to(dev)
                                for low in lows])
        return simulation.simulate_and_measure(
            tapes[0], n_qubits, type, obs, dens, shots=shots, key=key, initial_state=initial_state,
            batch=B, values=(angles.reshape(B, -1), mats),
        )

    def _execute_batched(self, type, obs, args, kwargs, in_axes, shots, key, initial_state):
        if not isinstance(in_axes, (tuple, list)):
            in_axes = (in_axes,) * len(args)
        if len(in_axes) != len(args):
            raise ValueError(f"in_axes has {len(in_axes)} entries but {len(args)} args were given")
        args = tuple(_to_tensor(a) for a in args)
        sizes = {a.shape[ax] for a, ax in zip(args, in_axes) if ax is not None}
        if len(sizes) != 1:
            raise ValueError(f"batched arguments must share one batch size, got {sorted(sizes)}")
        B = sizes.pop()
        if initial_state is not None and np.ndim(initial_state) == 2 and np.shape(initial_state)[0] != B:
            raise ValueError("2D initial_state must have one state per batch sample")
        box = {}
        dev = next((a.device for a, ax in zip(args, in_axes) if ax is not None), torch.device("cpu"))

        def single(*a):
            with recording() as tape, vectorized_recording():
                self.f(*a, **kwargs)
            box["tape"] = tape
            n_q = self._n_qubits or simulation.infer_n_qubits(tape, obs)
            # the values come from the same lowering the simulation uses, so
            # their order matches the plan for statevectors and density
            # matrices alike (channels contribute their superoperators)
            dens = simulation.has_noise(tape)
            low = lowering.lower_mixed(tape, n_q) if dens else lowering.lower_pure(tape, n_q)
            # constants join the batched values on the arguments' device
            vals = [v.to(dev, torch.float64) if isinstance(v, torch.Tensor)
                    else torch.as_tensor(float(v), dtype=torch.float64, device=dev) for v in low.angle_values]
            mats = [_as_tensor(v).to(dev, torch.complex128).reshape(-1)
                    for _, v in low.matrix_values]
            return (torch.stack(vals) if vals else torch.zeros(0, dtype=torch.float64), tuple(mats))

        try:
            angles, mats = torch.func.vmap(single, in_dims=tuple(in_axes))(*args)
        except RecordPerSample:
            return self._execute_per_sample(type, obs, args, kwargs, in_axes, B, shots, key, initial_state)
        tape = box["tape"]
        n_qubits = self._n_qubits or simulation.infer_n_qubits(tape, obs)
        use_density = simulation.has_noise(tape)
        mat = torch.cat([m.reshape(B, -1) for m in mats], dim=1) if len(mats) else None
        return simulation.simulate_and_measure(
            tape, n_qubits, type, obs, use_density, shots=shots, key=key, initial_state=initial_state,
            batch=B, values=(angles.reshape(B, -1), mat),
        )
