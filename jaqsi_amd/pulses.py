"""Pulse-level gates (mirrors jaqsi/pulses.py).

``Gates.RX(w, wires, pulse=True, pulse_params=...)`` records a
:class:`~jaqsi_amd.evolution.PendingEvolution` of the gate's drive
Hamiltonian; ``simulate_and_measure`` solves every pending gate of a tape in
one batched propagator launch (``jq_evolve``) and applies the unitaries as
explicit matrices, so gradients flow from the circuit's adjoint back into
the pulse parameters.

* :class:`PulseEnvelope` -- the envelope shapes and their registry of
  calibrated defaults (pulses.py:324-528);
* :meth:`PulseEnvelope.build_coeff_fns` -- the interaction-picture
  coefficient functions under the RWA, or exactly in the drive or lab frame
  (pulses.py:549-764);
* :class:`PulseParams` / :class:`PulseInformation` -- parameter trees of the
  leaf gates (RX, RY, RZ, CZ) and of the composite gates built from them
  (pulses.py:767-1141);
* :class:`PulseGates` -- the gates (pulses.py:1144-1857).

Envelopes and coefficient functions are written with torch operations, so
they evaluate on the device at every sample time of the solver in one call
and are differentiable in the pulse parameters and the pulse duration.
"""

from __future__ import annotations

import csv
import logging
import os
from contextlib import contextmanager
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Tuple, Union

import numpy as np
import torch

from jaqsi_amd.evolution import Evolution  # noqa: F401  (re-export)
from jaqsi_amd.operations import Hermitian
from jaqsi_amd.unitary import UnitaryGates

log = logging.getLogger(__name__)


def _t(x, like):
    return x if isinstance(x, torch.Tensor) else torch.as_tensor(x, dtype=torch.float64, device=like.device)


# ---------------------------------------------------------------------------
# envelopes (pulses.py:324-420)
# ---------------------------------------------------------------------------
class PulseEnvelope:
    """Registry of pulse envelope shapes ``(p, t, t_c) -> amplitude``; the
    carrier is applied by the coefficient functions."""

    @staticmethod
    def gaussian(p, t, t_c):
        """Lifted Gaussian ``A (g(t) - g(0)) / (1 - g(0))``, ``p = [A, sigma, ...]``."""
        return p[0] * PulseEnvelope._lifted_gaussian(p[1], t, t_c)

    @staticmethod
    def _lifted_gaussian(sigma, t, t_c):
        """Unit-peak lifted Gaussian.  With a = log g(t), b = log g(0) and
        d = a - b = t (2 t_c - t) / (2 sigma^2), the numerator
        exp(max(a, b)) (expm1(min(d, 0)) - expm1(min(-d, 0))) avoids the
        cancellation of g(t) - g(0) for wide pulses and stays finite outside
        [0, 2 t_c]."""
        t_c = _t(t_c, t)
        sigma = _t(sigma, t)
        a = -0.5 * ((t - t_c) / sigma) ** 2
        b = -0.5 * (t_c / sigma) ** 2
        d = t * (2 * t_c - t) / (2 * sigma**2)
        rise = torch.expm1(torch.clamp(d, max=0.0)) - torch.expm1(torch.clamp(-d, max=0.0))
        return torch.exp(torch.maximum(a, b * torch.ones_like(a))) * rise / -torch.expm1(b)

    @staticmethod
    def square(p, t, t_c):
        """Rectangle of height A and the given width, ``p = [A, width, ...]``."""
        return p[0] * (torch.abs(t - t_c) <= p[1] / 2).to(t.dtype)

    @staticmethod
    def cosine(p, t, t_c):
        """Raised cosine ``A cos(pi x)``, x = clip((t - t_c) / width, +-1/2)."""
        x = torch.clamp((t - t_c) / p[1], -0.5, 0.5)
        return p[0] * torch.cos(torch.pi * x)

    @staticmethod
    def drag(p, t, t_c):
        """DRAG in-phase envelope: the lifted Gaussian, ``p = [A, beta, sigma, ...]``."""
        return p[0] * PulseEnvelope._lifted_gaussian(p[2], t, t_c)

    @staticmethod
    def drag_quadrature(p, t, t_c):
        """DRAG quadrature ``-beta dE/dt`` of ``E = drag(p, t, t_c)``:
        A beta (t - t_c) / sigma^2 g(t) / (1 - g(0))."""
        A, beta, sigma = p[0], p[1], p[2]
        g = torch.exp(-0.5 * ((t - t_c) / sigma) ** 2)
        lift = -torch.expm1(-0.5 * (_t(t_c, t) / sigma) ** 2)
        return A * beta * (t - t_c) / sigma**2 * g / lift

    @staticmethod
    def _no_quadrature(p, t, t_c):
        return torch.zeros_like(t)

    @staticmethod
    def sech(p, t, t_c):
        """Hyperbolic secant ``A / cosh((t - t_c) / sigma)``."""
        return p[0] / torch.cosh((t - t_c) / p[1])

    @staticmethod
    def drag_legacy(p, t, t_c):
        """The single-quadrature ``g + beta dg`` model with the running centre
        t / 2 (``t_c`` ignored), ``p = [A, beta, sigma, ...]``."""
        A, beta, sigma = p[0], p[1], p[2]
        tc = t / 2
        g = A * torch.exp(-0.5 * ((t - tc) / sigma) ** 2)
        dg = g * (-(t - tc) / sigma**2)
        return g + beta * dg

    # Calibrated defaults of the leaf gates, [envelope params..., T]: the
    # reference's table (pulses.py:426-528), kept identical so that
    # pulse-level circuits give the same unitaries.
    REGISTRY = {
        "gaussian": {
            "fn": gaussian.__func__,
            "n_envelope_params": 2,
            "defaults": {
                "RX": np.array([0.44506904345408577, 1.9072729604580203, 3.5213966860043175]),
                "RY": np.array([0.449245182920846, 1.8896592598281816, 3.4886409030956145]),
            },
        },
        "square": {
            "fn": square.__func__,
            "n_envelope_params": 2,
            "defaults": {
                "RX": np.array([1.209660477974627, 0.8266782440135297, 1.1483171598013935]),
                "RY": np.array([1.0260908021252053, 1.000371727035551, 0.9745726190714447]),
            },
        },
        "cosine": {
            "fn": cosine.__func__,
            "n_envelope_params": 2,
            "defaults": {
                "RX": np.array([1.4970963752879989, 1.1647261280710832, 0.8317222416617329]),
                "RY": np.array([1.4970963752879989, 1.1647261280710832, 0.8317222416617329]),
            },
        },
        "drag": {
            "fn": drag.__func__,
            "quadrature_fn": drag_quadrature.__func__,
            "n_envelope_params": 3,
            "defaults": {
                "RX": np.array([0.4462644494309652, 5.56779671979968e-19, 1.890265883325973, 3.5141632250053174]),
                "RY": np.array([0.45045111802374943, 1.2007294980651857e-13, 1.872843283708579, 3.481473844222316]),
            },
        },
        "sech": {
            "fn": sech.__func__,
            "n_envelope_params": 2,
            "defaults": {
                "RX": np.array([1.18559023969428, 1.5022530879336038, 0.8547640992860961]),
                "RY": np.array([1.18559023969428, 1.5022530879336038, 0.8547640992860961]),
            },
        },
        "drag_legacy": {
            "fn": drag_legacy.__func__,
            "n_envelope_params": 3,
            "defaults": {
                "RX": np.array([0.326562746114197, 0.4002767596709071, 5.3228107728890315, 3.141300761986467]),
                "RY": np.array([0.323287924190616, 0.4065017233024265, 7.00299644871222, 3.139481229843545]),
            },
        },
        "general": {
            "fn": None,
            "n_envelope_params": 0,
            "defaults": {"RZ": np.array([0.5]), "CZ": np.array([0.3183098783513154])},
        },
    }

    @staticmethod
    def available() -> List[str]:
        return list(PulseEnvelope.REGISTRY.keys())

    @staticmethod
    def get(name: str) -> dict:
        if name not in PulseEnvelope.REGISTRY:
            raise ValueError(f"Unknown pulse envelope '{name}'. Available: {PulseEnvelope.available()}")
        return PulseEnvelope.REGISTRY[name]

    @staticmethod
    def build_coeff_fns(envelope_fn: Callable, omega_c: float, omega_q: float, rwa: bool = True,
                        frame: str = "drive", quadrature_fn: Optional[Callable] = None) -> Tuple[Callable, ...]:
        """The four interaction-picture coefficient functions
        (RX_X, RX_Y, RY_X, RY_Y) of pulses.py:549-764.

        The drive S(t) = E(t) cos(w_c t + phi) - Q(t) sin(w_c t + phi) on X,
        in the frame of w_q / 2 Z, gives S(t) [cos(w_q t) X - sin(w_q t) Y];
        RX drives with phi = 0 and RY with phi = pi / 2.  Under the RWA only
        the slow part survives: (E / 2)(cos phi X + sin phi Y) +
        (Q / 2)(-sin phi X + cos phi Y).  Without it, ``"lab"`` keeps the
        products as written and ``"drive"`` rewrites them with the
        product-to-sum identities around w_c -+ w_q.  The parameter vector is
        ``[envelope params..., T, w]``; the envelope is centred at T / 2 and
        every coefficient is scaled by the rotation angle w."""
        if frame not in ("lab", "drive"):
            raise ValueError(f"Unknown frame {frame!r}; expected 'lab' or 'drive'.")
        q_fn = quadrature_fn if quadrature_fn is not None else PulseEnvelope._no_quadrature
        if rwa:
            def rx_x(p, t):
                return 0.5 * envelope_fn(p, t, p[-2] / 2) * p[-1]

            def rx_y(p, t):
                return 0.5 * q_fn(p, t, p[-2] / 2) * p[-1]

            def ry_x(p, t):
                return -0.5 * q_fn(p, t, p[-2] / 2) * p[-1]

            def ry_y(p, t):
                return 0.5 * envelope_fn(p, t, p[-2] / 2) * p[-1]

            return rx_x, rx_y, ry_x, ry_y

        if frame == "drive":
            wd, ws = omega_c - omega_q, omega_c + omega_q

            def _eq(p, t):
                return envelope_fn(p, t, p[-2] / 2), q_fn(p, t, p[-2] / 2)

            def rx_x(p, t):
                E, Q = _eq(p, t)
                return (E * 0.5 * (torch.cos(wd * t) + torch.cos(ws * t))
                        - Q * 0.5 * (torch.sin(ws * t) + torch.sin(wd * t))) * p[-1]

            def rx_y(p, t):
                E, Q = _eq(p, t)
                return (-E * 0.5 * (torch.sin(ws * t) - torch.sin(wd * t))
                        + Q * 0.5 * (torch.cos(wd * t) - torch.cos(ws * t))) * p[-1]

            def ry_x(p, t):
                E, Q = _eq(p, t)
                return (-E * 0.5 * (torch.sin(ws * t) + torch.sin(wd * t))
                        - Q * 0.5 * (torch.cos(wd * t) + torch.cos(ws * t))) * p[-1]

            def ry_y(p, t):
                E, Q = _eq(p, t)
                return (-E * 0.5 * (torch.cos(ws * t) - torch.cos(wd * t))
                        + Q * 0.5 * (torch.sin(ws * t) - torch.sin(wd * t))) * p[-1]

            return rx_x, rx_y, ry_x, ry_y

        def _drive(p, t, phase):
            E, Q = envelope_fn(p, t, p[-2] / 2), q_fn(p, t, p[-2] / 2)
            return E * torch.cos(omega_c * t + phase) - Q * torch.sin(omega_c * t + phase)

        def rx_x(p, t):
            return _drive(p, t, 0.0) * torch.cos(omega_q * t) * p[-1]

        def rx_y(p, t):
            return -_drive(p, t, 0.0) * torch.sin(omega_q * t) * p[-1]

        def ry_x(p, t):
            return _drive(p, t, np.pi / 2) * torch.cos(omega_q * t) * p[-1]

        def ry_y(p, t):
            return -_drive(p, t, np.pi / 2) * torch.sin(omega_q * t) * p[-1]

        return rx_x, rx_y, ry_x, ry_y


# ---------------------------------------------------------------------------
# parameter trees (pulses.py:21-304)
# ---------------------------------------------------------------------------
@dataclass
class DecompositionStep:
    """One step of a composite gate: child gate, wire selector ("all",
    "target", "control") and the child's angle as a function of the parent's."""

    gate: "PulseParams"
    wire_fn: str = "all"
    angle_fn: Optional[Callable] = None


@dataclass(frozen=True)
class PulseStateSnapshot:
    envelope: str
    rwa: bool
    frame: str
    leaf_params: Dict[str, np.ndarray]


def _cat(parts):
    if any(isinstance(x, torch.Tensor) for x in parts):
        return torch.cat([torch.as_tensor(x, dtype=torch.float64).reshape(-1) if not isinstance(x, torch.Tensor)
                          else x.reshape(-1).to(torch.float64) for x in parts])
    return np.concatenate([np.ravel(np.asarray(x, dtype=np.float64)) for x in parts])


class PulseParams:
    """Leaf nodes hold parameters, composite nodes a decomposition into
    child gates; ``params`` of a composite is the concatenation of its
    children's (pulses.py:47-304)."""

    def __init__(self, name: str = "", params=None, decomposition: Optional[List[DecompositionStep]] = None) -> None:
        assert (params is None) != (decomposition is None), "Exactly one of `params` or `decomposition` must be provided."
        self.decomposition = decomposition
        self._pulse_obj = [s.gate for s in decomposition] if decomposition else None
        if params is not None:
            self._params = params
        self.name = name

    def __len__(self) -> int:
        if self.is_leaf:
            return len(self._params)
        return sum(len(c) for c in self.childs)

    def __getitem__(self, idx):
        return self._params[idx] if self.is_leaf else self.childs[idx].params

    def __str__(self) -> str:
        return self.name

    __repr__ = __str__

    @property
    def is_leaf(self) -> bool:
        return self._pulse_obj is None

    @property
    def size(self) -> int:
        return len(self)

    @property
    def leafs(self) -> List["PulseParams"]:
        if self.is_leaf:
            return [self]
        out = []
        for c in self.childs:
            for lf in c.leafs:
                if lf not in out:
                    out.append(lf)
        return out

    @property
    def childs(self) -> List["PulseParams"]:
        return [] if self.is_leaf else self._pulse_obj

    @property
    def shape(self):
        if self.is_leaf:
            return [len(self._params)]
        return [c.shape for c in self.childs]

    @property
    def params(self):
        if self.is_leaf:
            return self._params
        return _cat(self.split_params(None, leafs=False))

    @params.setter
    def params(self, value) -> None:
        if self.is_leaf:
            self._params = value
            return
        idx = 0
        for c in self.childs:
            c.params = value[idx: idx + c.size]
            idx += c.size

    @property
    def leaf_params(self):
        if self.is_leaf:
            return self._params
        return _cat(self.split_params(None, leafs=True))

    @leaf_params.setter
    def leaf_params(self, value) -> None:
        if self.is_leaf:
            self._params = value
            return
        idx = 0
        for lf in self.leafs:
            lf.params = value[idx: idx + lf.size]
            idx += lf.size

    def split_params(self, params=None, leafs: bool = False):
        if params is None:
            if self.is_leaf:
                return self._params
            return [o.params for o in (self.leafs if leafs else self.childs)]
        if self.is_leaf:
            return params
        out, idx = [], 0
        for o in (self.leafs if leafs else self.childs):
            out.append(params[idx: idx + o.size])
            idx += o.size
        return out


class PulseParamManager:
    """Hands consecutive slices of one model pulse-parameter vector to the
    gates recorded under ``Gates.pulse_manager_context`` (pulses.py:1860-1872)."""

    def __init__(self, pulse_params):
        self.pulse_params = pulse_params
        self.idx = 0

    def get(self, n: int):
        if self.idx + n > len(self.pulse_params):
            raise ValueError("Not enough pulse parameters left for this gate")
        params = self.pulse_params[self.idx: self.idx + n]
        params = params.squeeze() if isinstance(params, torch.Tensor) else np.squeeze(params)
        self.idx += n
        return params


class PulseInformation:
    """Active envelope, RWA flag and frame, and the parameter trees of every
    pulse gate (pulses.py:767-1141)."""

    DEFAULT_ENVELOPE: str = "gaussian"
    DEFAULT_RWA: bool = True
    DEFAULT_FRAME: str = "drive"
    LEAF_GATE_NAMES: Tuple[str, ...] = ("RX", "RY", "RZ", "CZ")
    OPTIMIZED_PULSES: Dict[str, np.ndarray] = {}

    _envelope: str = DEFAULT_ENVELOPE
    _rwa: bool = DEFAULT_RWA
    _frame: str = DEFAULT_FRAME

    @classmethod
    def _build_leaf_gates(cls):
        d = PulseEnvelope.get(cls._envelope)["defaults"]
        g = PulseEnvelope.get("general")["defaults"]
        cls.RX = PulseParams(name="RX", params=np.array(d["RX"], dtype=np.float64))
        cls.RY = PulseParams(name="RY", params=np.array(d["RY"], dtype=np.float64))
        cls.RZ = PulseParams(name="RZ", params=np.array(g["RZ"], dtype=np.float64))
        cls.CZ = PulseParams(name="CZ", params=np.array(g["CZ"], dtype=np.float64))

    @classmethod
    def _build_composite_gates(cls):
        S = DecompositionStep
        pi = np.pi
        cls.H = PulseParams(name="H", decomposition=[S(cls.RZ, "all", lambda w: pi), S(cls.RY, "all", lambda w: pi / 2)])
        cls.CX = PulseParams(name="CX", decomposition=[S(cls.H, "target", lambda w: 0.0), S(cls.CZ, "all", lambda w: 0.0),
                                                       S(cls.H, "target", lambda w: 0.0)])
        cls.CY = PulseParams(name="CY", decomposition=[S(cls.RZ, "target", lambda w: -pi / 2), S(cls.CX, "all"),
                                                       S(cls.RZ, "target", lambda w: pi / 2)])
        cls.CRX = PulseParams(name="CRX", decomposition=[
            S(cls.RZ, "target", lambda w: pi / 2), S(cls.RY, "target", lambda w: w / 2), S(cls.CX, "all", lambda w: 0.0),
            S(cls.RY, "target", lambda w: -w / 2), S(cls.CX, "all", lambda w: 0.0), S(cls.RZ, "target", lambda w: -pi / 2)])
        cls.CRY = PulseParams(name="CRY", decomposition=[
            S(cls.RY, "target", lambda w: w / 2), S(cls.CX, "all", lambda w: 0.0), S(cls.RY, "target", lambda w: -w / 2),
            S(cls.CX, "all", lambda w: 0.0)])
        cls.CRZ = PulseParams(name="CRZ", decomposition=[
            S(cls.RZ, "target", lambda w: w / 2), S(cls.CX, "all", lambda w: 0.0), S(cls.RZ, "target", lambda w: -w / 2),
            S(cls.CX, "all", lambda w: 0.0)])
        cls.CPhase = PulseParams(name="CPhase", decomposition=[
            S(cls.RZ, "control", lambda w: w / 2), S(cls.RZ, "target", lambda w: w / 2), S(cls.CX, "all", lambda w: 0.0),
            S(cls.RZ, "target", lambda w: -w / 2), S(cls.CX, "all", lambda w: 0.0)])
        cls.RZZ = PulseParams(name="RZZ", decomposition=[
            S(cls.CX, "all", lambda w: 0.0), S(cls.RZ, "target", lambda w: w), S(cls.CX, "all", lambda w: 0.0)])
        cls.RXX = PulseParams(name="RXX", decomposition=[
            S(cls.H, "control", lambda w: 0.0), S(cls.H, "target", lambda w: 0.0), S(cls.CX, "all", lambda w: 0.0),
            S(cls.RZ, "target", lambda w: w), S(cls.CX, "all", lambda w: 0.0), S(cls.H, "control", lambda w: 0.0),
            S(cls.H, "target", lambda w: 0.0)])
        cls.RYY = PulseParams(name="RYY", decomposition=[
            S(cls.RX, "control", lambda w: pi / 2), S(cls.RX, "target", lambda w: pi / 2), S(cls.CX, "all", lambda w: 0.0),
            S(cls.RZ, "target", lambda w: w), S(cls.CX, "all", lambda w: 0.0), S(cls.RX, "control", lambda w: -pi / 2),
            S(cls.RX, "target", lambda w: -pi / 2)])
        cls.RZX = PulseParams(name="RZX", decomposition=[
            S(cls.H, "target", lambda w: 0.0), S(cls.CX, "all", lambda w: 0.0), S(cls.RZ, "target", lambda w: w),
            S(cls.CX, "all", lambda w: 0.0), S(cls.H, "target", lambda w: 0.0)])
        cls.Rot = PulseParams(name="Rot", decomposition=[
            S(cls.RZ, "all", lambda w: w[0]), S(cls.RY, "all", lambda w: w[1]), S(cls.RZ, "all", lambda w: w[2])])
        cls.unique_gate_set = [cls.RX, cls.RY, cls.RZ, cls.CZ]

    @classmethod
    def set_envelope(cls, name: str, rwa: Optional[bool] = None, frame: Optional[str] = None) -> None:
        """Switch the envelope (and optionally the RWA flag / frame), rebuild
        the parameter trees and the coefficient functions of PulseGates."""
        info = PulseEnvelope.get(name)
        cls._envelope = name
        if rwa is not None:
            cls._rwa = bool(rwa)
        if frame is not None:
            if frame not in ("lab", "drive"):
                raise ValueError(f"Unknown frame {frame!r}; expected 'lab' or 'drive'.")
            cls._frame = frame
        cls._build_leaf_gates()
        cls._build_composite_gates()
        fns = PulseEnvelope.build_coeff_fns(info["fn"], PulseGates.omega_c, PulseGates.omega_q, rwa=cls._rwa,
                                            frame=cls._frame, quadrature_fn=info.get("quadrature_fn"))
        PulseGates._coeff_RX_X, PulseGates._coeff_RX_Y, PulseGates._coeff_RY_X, PulseGates._coeff_RY_Y = (
            staticmethod(f) for f in fns)
        PulseGates._coeff_Sx = staticmethod(fns[0])
        PulseGates._coeff_Sy = staticmethod(fns[3])
        PulseGates._active_envelope = name
        PulseGates._active_quadrature = "quadrature_fn" in info
        PulseGates._active_rwa = cls._rwa
        PulseGates._active_frame = cls._frame

    @classmethod
    def set_rwa(cls, rwa: bool) -> None:
        cls.set_envelope(cls._envelope, rwa=bool(rwa))

    @classmethod
    def get_envelope(cls) -> str:
        return cls._envelope

    @classmethod
    def get_rwa(cls) -> bool:
        return cls._rwa

    @classmethod
    def set_frame(cls, frame: str) -> None:
        cls.set_envelope(cls._envelope, frame=str(frame))

    @classmethod
    def get_frame(cls) -> str:
        return cls._frame

    @classmethod
    def snapshot_state(cls) -> PulseStateSnapshot:
        leaf = {n: np.array(_host(getattr(cls, n).params)) for n in cls.LEAF_GATE_NAMES}
        return PulseStateSnapshot(cls._envelope, cls._rwa, cls._frame, leaf)

    @classmethod
    def restore_state(cls, snap: PulseStateSnapshot) -> None:
        cls.set_envelope(snap.envelope, rwa=snap.rwa, frame=snap.frame)
        for name, params in snap.leaf_params.items():
            gate = cls.gate_by_name(name)
            if gate is None or not gate.is_leaf:
                raise ValueError(f"Cannot restore unknown leaf pulse gate {name!r}.")
            if np.shape(_host(gate.params)) != np.shape(params):
                raise ValueError(f"Snapshot for {name!r} has shape {np.shape(params)}, but active gate expects "
                                 f"{np.shape(_host(gate.params))}.")
            gate.params = params

    @classmethod
    @contextmanager
    def preserve_state(cls):
        snap = cls.snapshot_state()
        try:
            yield snap
        finally:
            cls.restore_state(snap)

    @classmethod
    def reset_defaults(cls, envelope: Optional[str] = None, rwa: Optional[bool] = None,
                       frame: Optional[str] = None) -> None:
        cls.set_envelope(cls.DEFAULT_ENVELOPE if envelope is None else envelope,
                         rwa=cls.DEFAULT_RWA if rwa is None else rwa,
                         frame=cls.DEFAULT_FRAME if frame is None else frame)

    @staticmethod
    def gate_by_name(gate):
        return getattr(PulseInformation, gate if isinstance(gate, str) else gate.__name__, None)

    @staticmethod
    def num_params(gate) -> int:
        return len(PulseInformation.gate_by_name(gate))

    @classmethod
    def update_params(cls, path: Optional[str] = None) -> None:
        """Load optimised pulse parameters from a QOC results CSV (rows
        ``gate, fidelity, params...``) into OPTIMIZED_PULSES
        (pulses.py:1104-1131).  ``path`` defaults to
        ``qoc_results_<envelope>.csv`` next to this module."""
        if path is None:
            path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"qoc_results_{cls.get_envelope()}.csv")
        if not os.path.isfile(path):
            log.error(f"No optimized pulses found at {path}")
            return
        with open(path, "r") as fh:
            for row in csv.reader(fh):
                cls.OPTIMIZED_PULSES[row[0]] = np.array([float(x) for x in row[2:]])


def _host(x):
    return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


def _pack_params(pulse_params, keep: slice, w):
    """ravel(pulse_params)[keep] followed by w: numpy while every value is a
    plain number (so that identical gates are recognised and solved once),
    a torch tensor as soon as one is a tensor (so gradients flow)."""
    if isinstance(pulse_params, torch.Tensor) or isinstance(w, torch.Tensor):
        pp = pulse_params if isinstance(pulse_params, torch.Tensor) else torch.as_tensor(np.asarray(pulse_params, dtype=np.float64))
        ww = w if isinstance(w, torch.Tensor) else torch.as_tensor(np.asarray(w, dtype=np.float64))
        dev = pp.device if pp.device.type != "cpu" else ww.device
        return torch.cat([pp.reshape(-1).to(dev, torch.float64)[keep], ww.reshape(-1).to(dev, torch.float64)])
    return np.concatenate([np.ravel(np.asarray(pulse_params, dtype=np.float64))[keep], np.ravel(np.asarray(w, dtype=np.float64))])


def _duration(pulse_params):
    return pulse_params[-1]


# ---------------------------------------------------------------------------
# the gates (pulses.py:1144-1857)
# ---------------------------------------------------------------------------
class PulseGates:
    """Pulse-level implementations of the gates: RX / RY are driven
    rotations of the active envelope, RZ a virtual Z rotation, CZ a ZZ
    coupling pulse; the other gates are composites of these (following
    https://doi.org/10.5445/IR/1000184129, as the reference does)."""

    omega_q = 10 * np.pi
    omega_c = 10 * np.pi

    X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
    Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
    Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
    Id = np.eye(2, dtype=np.complex128)
    _H_CZ = (np.pi / 4) * (np.kron(Id, Id) - np.kron(Z, Id) - np.kron(Id, Z) + np.kron(Z, Z))
    _H_corr = np.pi / 2 * np.eye(2, dtype=np.complex128)

    _active_envelope: str = "gaussian"
    _active_quadrature: bool = False
    _active_rwa: bool = True
    _active_frame: str = "drive"

    @staticmethod
    def _coeff_Sz(p, t):
        """RZ: p[0] * w over unit time."""
        return p[0] * p[1]

    @staticmethod
    def _coeff_Sc(p, t):
        """The H gate's global-phase correction."""
        return -1.0

    @staticmethod
    def _coeff_Scz(p, t):
        """CZ: p * pi over unit time."""
        return p * np.pi

    @staticmethod
    def _noise(wires, noise_params, random_key, w=None):
        if noise_params is not None and noise_params.get("GateError") is not None and w is not None:
            w, random_key = UnitaryGates.GateError(w, noise_params, random_key)
        return w

    @staticmethod
    def Rot(phi, theta, omega, wires, pulse_params=None, noise_params=None, random_key=None) -> None:
        if noise_params is not None and "GateError" in noise_params:
            phi, random_key = UnitaryGates.GateError(phi, noise_params, random_key)
            theta, random_key = UnitaryGates.GateError(theta, noise_params, random_key)
            omega, random_key = UnitaryGates.GateError(omega, noise_params, random_key)
        PulseGates._execute_composite("Rot", [phi, theta, omega], wires, pulse_params)
        UnitaryGates.Noise(wires, noise_params)

    @staticmethod
    def PauliRot(pauli, theta, wires, pulse_params=None, noise_params=None, random_key=None) -> None:
        raise NotImplementedError("PauliRot gate is not implemented as PulseGate")

    @staticmethod
    def RX(w, wires, pulse_params=None, noise_params=None, random_key=None) -> None:
        """X rotation driven by the active envelope: H = f_X X (+ f_Y Y for a
        second quadrature or without the RWA), params [envelope..., T]."""
        pulse_params = PulseInformation.RX.split_params(pulse_params)
        t = _duration(pulse_params)
        H_eff = PulseGates._coeff_RX_X * Hermitian(PulseGates.X, wires=wires, record=False)
        if not PulseGates._active_rwa or PulseGates._active_quadrature:
            H_eff = H_eff + PulseGates._coeff_RX_Y * Hermitian(PulseGates.Y, wires=wires, record=False)
        w = PulseGates._noise(wires, noise_params, random_key, w)
        env = _pack_params(pulse_params, slice(None), w)
        H_eff.evolve(name="RX")([env] * H_eff.n_terms, t)
        UnitaryGates.Noise(wires, noise_params)

    @staticmethod
    def RY(w, wires, pulse_params=None, noise_params=None, random_key=None) -> None:
        """Y rotation: carrier phase pi / 2, so the RWA part drives +Y."""
        pulse_params = PulseInformation.RY.split_params(pulse_params)
        t = _duration(pulse_params)
        H_eff = PulseGates._coeff_RY_Y * Hermitian(PulseGates.Y, wires=wires, record=False)
        if not PulseGates._active_rwa or PulseGates._active_quadrature:
            H_eff = PulseGates._coeff_RY_X * Hermitian(PulseGates.X, wires=wires, record=False) + H_eff
        w = PulseGates._noise(wires, noise_params, random_key, w)
        env = _pack_params(pulse_params, slice(None), w)
        H_eff.evolve(name="RY")([env] * H_eff.n_terms, t)
        UnitaryGates.Noise(wires, noise_params)

    @staticmethod
    def RZ(w, wires, pulse_params=None, noise_params=None, random_key=None) -> None:
        """Virtual Z rotation exp(-i p w Z) over unit time (p = 1/2 by default)."""
        pulse_params = PulseInformation.RZ.split_params(pulse_params)
        H_eff = PulseGates._coeff_Sz * Hermitian(PulseGates.Z, wires=wires, record=False)
        w = PulseGates._noise(wires, noise_params, random_key, w)
        H_eff.evolve(name="RZ")([_pack_params(pulse_params, slice(None, 1), w)], 1)
        UnitaryGates.Noise(wires, noise_params)

    @staticmethod
    def _resolve_wires(wire_fn, wires):
        wl = [wires] if isinstance(wires, int) else list(wires)
        if wire_fn == "all":
            return wires if len(wl) > 1 else wl[0]
        if wire_fn == "target":
            return wl[-1]
        if wire_fn == "control":
            return wl[0]
        raise ValueError(f"Unknown wire_fn: {wire_fn!r}")

    @staticmethod
    def _execute_composite(gate_name, w, wires, pulse_params=None):
        """Walk the decomposition of a composite gate (pulses.py:1516-1568)."""
        pp = PulseInformation.gate_by_name(gate_name)
        parts = [None] * len(pp.decomposition) if pulse_params is None else pp.split_params(pulse_params)
        for step, child in zip(pp.decomposition, parts):
            cw = PulseGates._resolve_wires(step.wire_fn, wires)
            cang = step.angle_fn(w) if step.angle_fn is not None else w
            gate = getattr(PulseGates, step.gate.name)
            if step.gate.name in ("RX", "RY", "RZ"):
                gate(cang, wires=cw, pulse_params=child)
            elif step.gate.name == "CZ":
                gate(wires=cw, pulse_params=child)
            elif step.gate.name == "Rot":
                gate(*cang, wires=cw, pulse_params=child)
            elif step.gate.name in ("CRX", "CRY", "CRZ", "CPhase", "RXX", "RYY", "RZZ", "RZX"):
                gate(cang, wires=cw, pulse_params=child)
            else:
                gate(wires=cw, pulse_params=child)

    @staticmethod
    def H(wires, pulse_params=None, noise_params=None, random_key=None) -> None:
        """RZ(pi) RY(pi/2), then the correction phase exp(+i pi/2)."""
        PulseGates._execute_composite("H", 0.0, wires, pulse_params)
        H_corr = PulseGates._coeff_Sc * Hermitian(PulseGates._H_corr, wires=wires, record=False)
        H_corr.evolve(name="H")([0], 1)
        UnitaryGates.Noise(wires, noise_params)

    @staticmethod
    def CX(wires, pulse_params=None, noise_params=None, random_key=None) -> None:
        PulseGates._execute_composite("CX", 0.0, wires, pulse_params)
        UnitaryGates.Noise(wires, noise_params)

    @staticmethod
    def CY(wires, pulse_params=None, noise_params=None, random_key=None) -> None:
        PulseGates._execute_composite("CY", 0.0, wires, pulse_params)
        UnitaryGates.Noise(wires, noise_params)

    @staticmethod
    def CZ(wires, pulse_params=None, noise_params=None, random_key=None) -> None:
        """ZZ coupling pulse exp(-i p pi H_CZ) over unit time."""
        if pulse_params is None:
            pulse_params = PulseInformation.CZ.params
        H_eff = PulseGates._coeff_Scz * Hermitian(PulseGates._H_CZ, wires=wires, record=False)
        H_eff.evolve(name="CZ")([pulse_params], 1)
        UnitaryGates.Noise(wires, noise_params)


def _composite(name: str, with_angle: bool):
    if with_angle:
        def gate(w, wires, pulse_params=None, noise_params=None, random_key=None) -> None:
            w = PulseGates._noise(wires, noise_params, random_key, w)
            PulseGates._execute_composite(name, w, wires, pulse_params)
            UnitaryGates.Noise(wires, noise_params)
    else:
        def gate(wires, pulse_params=None, noise_params=None, random_key=None) -> None:
            PulseGates._execute_composite(name, 0.0, wires, pulse_params)
            UnitaryGates.Noise(wires, noise_params)
    gate.__name__ = name
    gate.__doc__ = f"{name} as the pulse decomposition of PulseInformation.{name}."
    return staticmethod(gate)


for _name in ("CRX", "CRY", "CRZ", "CPhase", "RXX", "RYY", "RZZ", "RZX"):
    setattr(PulseGates, _name, _composite(_name, True))

PulseInformation.reset_defaults()
