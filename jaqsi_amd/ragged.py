"""Ragged batches: circuits of several structures (e.g. uneven depths) with
shared observables, run as ONE forward + adjoint sweep through
``jq_sv_adjoint_groups``.

The reference can only vmap circuits of identical structure
(script.py:406-434), so a batch of uneven depths costs one execute() per
depth group there.  Here every pass step is a single launch over all
circuits that still have a pass at that step.
"""

from __future__ import annotations

from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from jaqsi_amd import _native as nat
from jaqsi_amd import lowering
from jaqsi_amd.simulation import _cdtype_torch, _dtype_code, _stream, device, get_plan
from jaqsi_amd.tape import recording


class RaggedBatch:
    """Groups of circuits; group g has ``batch[g]`` circuits of structure
    ``tapes[g]`` (a recorded tape) on ``n_qubits`` wires.  Observables are
    single-wire diagonals given as (wire, d0, d1) (PauliZ: (q, 1, -1)).

    Only expectation values and gradients leave the batch, so by default the
    circuits' trailing permutation / diagonal gates are measured through the
    observables (``fold_tail``) when every group ends in the same such tail
    (the groups share one observable table)."""

    def __init__(self, tapes: Sequence[list], batch: Sequence[int], n_qubits: int,
                 diags: Sequence[Tuple[int, float, float]], dtype: Optional[int] = None,
                 knobs: Optional[dict] = None, fold: bool = True):
        self.n_qubits = n_qubits
        self.dtype = _dtype_code() if dtype is None else dtype
        self.batch = list(batch)
        self.lowered = [lowering.lower_pure(t, n_qubits) for t in tapes]
        # one set of tuning knobs for every group (given, or the autotuner's
        # cached choice for each group's own structure)
        self.plans = [get_plan(low, self.dtype, batch=B, knobs=knobs, fold=fold)
                      for low, B in zip(self.lowered, self.batch)]
        if fold:
            folds = [p.fold() for p in self.plans]
            same = all(f[0] == folds[0][0] and (f[0] == 0 or (np.array_equal(f[1], folds[0][1]) and f[2] == folds[0][2]))
                       for f in folds)
            if not same:
                self.plans = [get_plan(low, self.dtype, batch=B, knobs=knobs)
                              for low, B in zip(self.lowered, self.batch)]
        self.n_params = [len(low.angle_values) for low in self.lowered]
        self.diags = list(diags)
        self.wires = np.array([d[0] for d in diags], dtype=np.int32)
        self.d = np.array([[d[1], d[2]] for d in diags], dtype=np.float64).reshape(-1)
        dev = device()
        self.mats = []
        for low, B in zip(self.lowered, self.batch):
            m = lowering.matrix_tensor(low.matrix_values, low.n_matrix_elems, None)
            self.mats.append(m.reshape(1, -1).expand(B, -1).contiguous().to(dev) if m is not None else None)
        self.state = [torch.empty((B, 1 << n_qubits), dtype=_cdtype_torch(self.dtype), device=dev) for B in self.batch]

    def specialise(self, mode: int = nat.JQ_JIT_SYNC) -> List[int]:
        """Plan-specialised kernels for every group's plan (nat.JQ_JIT_*).
        Once all groups have them loaded, ``adjoint`` runs each group through
        them; until then all groups share the interpreter's launches.
        Returns the compile phase of each plan.  Plans compile one after
        the other, so a group reuses the kernels of passes it shares with
        the groups before it (the run-time compiler does not overlap
        compiles)."""
        return [p.jit(mode) for p in self.plans]

    def angle_shapes(self) -> List[Tuple[int, int]]:
        return [(B, P) for B, P in zip(self.batch, self.n_params)]

    def adjoint(self, angles: Sequence[torch.Tensor], ct: Sequence[torch.Tensor]):
        """Per group: angles (B, P) float64, ct (B, n_obs) float64, all on the
        device.  Returns (expvals, grads) lists of (B, n_obs) and (B, P)."""
        dev = device()
        n_obs = len(self.diags)
        groups = (nat.Group * len(self.plans))()
        evs, grads, keep = [], [], []
        for g, plan in enumerate(self.plans):
            B = self.batch[g]
            a = angles[g]
            c = ct[g]
            if a.shape != (B, self.n_params[g]) or a.dtype != torch.float64 or not a.is_contiguous():
                raise ValueError(f"group {g}: angles must be a contiguous float64 ({B}, {self.n_params[g]}) tensor")
            if c.shape != (B, n_obs) or c.dtype != torch.float64 or not c.is_contiguous():
                raise ValueError(f"group {g}: ct must be a contiguous float64 ({B}, {n_obs}) tensor")
            ev = torch.empty((B, n_obs), dtype=torch.float64, device=dev)
            ga = torch.empty((B, max(1, self.n_params[g])), dtype=torch.float64, device=dev)
            gr = groups[g]
            gr.plan = plan.handle
            gr.batch = B
            gr.psi0_batched = 0
            gr.angles = nat.ptr(a)
            gr.matrices = nat.ptr(self.mats[g])
            gr.psi0 = None
            gr.ct = nat.ptr(c)
            gr.expvals = nat.ptr(ev)
            gr.grad_angles = nat.ptr(ga)
            gr.state = nat.ptr(self.state[g])
            evs.append(ev)
            grads.append(ga[:, : self.n_params[g]])
            keep.append((a, c))
        nat.check(nat.lib().jq_sv_adjoint_groups(len(self.plans), groups, nat.JQ_OBS_SV_DIAG, n_obs,
                                                 nat.ptr(self.wires), nat.ptr(self.d), _stream()))
        return evs, grads


def record(fn, *args) -> list:
    """Record one circuit structure (the values recorded are irrelevant:
    plans depend on the structure only)."""
    with recording() as tape:
        fn(*args)
    return tape
