"""Autotuner of the plan knobs: the low global bits every tile holds, the
tile size and the kernel variants (``low_bits`` / ``tile_bits`` /
``variant_flags`` of ``jq_plan_desc``).

For one circuit structure, dtype and batch it times a forward + adjoint
step of each candidate plan on the GPU, by coordinate descent (first
``low_bits`` with the built-in variants, then, for complex64 circuits of 14
or more wires, 14-bit tiles on the best ``low_bits``, then each variant
flag on the best of those), and keeps the fastest.  Choices are cached per GPU
architecture, dtype, batch bucket (the next power of two) and structure
hash:

* a user database (``$JAQSI_AMD_TUNE_DB``, else
  ``~/.cache/jaqsi_amd/tuning.json``), written by ``tune``;
* a table shipped with the package (``jaqsi_amd/tuning/<arch>.json``),
  produced by ``python -m jaqsi_amd.autotune`` on that GPU.

``config_for`` is what plan creation consults: in mode ``"lookup"`` (the
default) it returns a cached choice or the built-in knobs; in ``"tune"`` it
tunes on a cache miss; ``"off"`` always gives the built-in knobs.  No
reference counterpart: the reference leaves tiling to XLA.
"""

from __future__ import annotations

import argparse
import hashlib
import json
import os
import statistics
import threading
import time
from typing import Dict, List, Optional

import torch

from jaqsi_amd import _native as nat

DEFAULT = {"low_bits": 0, "variant_flags": 0, "tile_bits": 0}
_LOW_CANDIDATES = {0: (5, 4, 6, 3), 1: (6, 5, 4)}  # complex64 / complex128, built-in first
_FLAG_CANDIDATES = {0: (nat.JQ_TUNE_BWD16, nat.JQ_TUNE_NO_PERSIST_BWD, nat.JQ_TUNE_NO_DMA_FWD, nat.JQ_TUNE_NARROW_BWD),
                    1: (nat.JQ_TUNE_NO_PERSIST_BWD,)}
_mode = os.environ.get("JAQSI_AMD_AUTOTUNE", "lookup")
_MIN_GAIN = 0.01  # smallest relative gain that moves a knob off its current choice
_lock = threading.Lock()
_db: Optional[dict] = None
_shipped: Dict[str, dict] = {}
_memo: Dict[tuple, dict] = {}  # (structure key, dtype, bucket, density, mode) -> knobs


def set_mode(mode: str) -> None:
    """"off", "lookup" (cached choices only) or "tune" (tune on a miss)."""
    global _mode
    if mode not in ("off", "lookup", "tune"):
        raise ValueError(f"unknown autotune mode {mode!r}")
    _mode = mode
    _memo.clear()


def get_mode() -> str:
    return _mode


def arch() -> str:
    return torch.cuda.get_device_properties(torch.cuda.current_device()).gcnArchName.split(":")[0]


def structure_hash(key: tuple) -> str:
    return hashlib.sha1(repr(key).encode()).hexdigest()[:20]


def batch_bucket(batch: int) -> int:
    return 1 << max(0, int(batch) - 1).bit_length()


def entry_key(arch_name: str, low_key: tuple, dtype: int, batch: int, fold: bool = False) -> str:
    """Cache key; ``fold`` marks plans whose trailing permutation / diagonal
    gates are measured through the observables (fewer passes)."""
    k = f"{arch_name}|c{128 if dtype else 64}|b{batch_bucket(batch)}|{structure_hash(low_key)}"
    return k + "|fold" if fold else k


def db_path() -> str:
    return os.environ.get("JAQSI_AMD_TUNE_DB", os.path.join(os.path.expanduser("~"), ".cache", "jaqsi_amd",
                                                            "tuning.json"))


def _load_db() -> dict:
    global _db
    if _db is None:
        try:
            with open(db_path()) as f:
                _db = json.load(f)
        except (OSError, ValueError):
            _db = {}
    return _db


def _shipped_table(arch_name: str) -> dict:
    if arch_name not in _shipped:
        p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", f"{arch_name}.json")
        try:
            with open(p) as f:
                _shipped[arch_name] = json.load(f)
        except (OSError, ValueError):
            _shipped[arch_name] = {}
    return _shipped[arch_name]


def lookup(low_key: tuple, dtype: int, batch: int, fold: bool = False) -> Optional[dict]:
    """The cached choice (user database first, then the shipped table); a
    folded plan without an entry of its own takes the unfolded one's."""
    a = arch()
    keys = [entry_key(a, low_key, dtype, batch, True)] if fold else []
    keys.append(entry_key(a, low_key, dtype, batch))
    for k in keys:
        with _lock:
            rec = _load_db().get(k)
        if rec is None:
            rec = _shipped_table(a).get(k)
            if rec is not None:
                rec = dict(rec, source="table")
        elif "source" not in rec:
            rec = dict(rec, source="database")
        if rec is not None:
            return rec
    return None


def _store(k: str, rec: dict) -> None:
    with _lock:
        db = _load_db()
        db[k] = rec
        p = db_path()
        try:
            os.makedirs(os.path.dirname(p), exist_ok=True)
            tmp = p + f".{os.getpid()}.tmp"
            with open(tmp, "w") as f:
                json.dump(db, f, indent=1, sort_keys=True)
            os.replace(tmp, p)
        except OSError:
            pass  # a read-only home keeps the choice for this process only


def _time_step(plan, batch: int, n_params: int, mats, density: bool, reps: int, rounds: int) -> float:
    """Median over rounds of the mean time (ms) of one forward + adjoint."""
    from jaqsi_amd.simulation import _Runner

    dev = torch.device("cuda", torch.cuda.current_device())
    g = torch.Generator(device="cpu").manual_seed(7)
    angles = (torch.rand((batch, n_params), generator=g, dtype=torch.float64) * 6.283).to(dev)
    m = mats.reshape(1, -1).expand(batch, -1).contiguous().to(dev) if mats is not None else None
    n_q = plan.n_wires // 2 if density else plan.n_wires
    runner = _Runner(plan, batch, density, [(0, 1.0, -1.0)], n_q, False, plan.dtype)
    ct = torch.ones((batch, 1), dtype=torch.float64, device=dev)

    def step():
        runner.adjoint(angles, m, None, ct, False)

    step()
    torch.cuda.synchronize()
    times = []
    for _ in range(rounds):
        t0 = time.perf_counter()
        for _ in range(reps):
            step()
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) / reps * 1e3)
    return statistics.median(times)


def tune(low, dtype: int, batch: int, density: bool = False, jit: Optional[bool] = None, reps: int = 3,
         rounds: int = 3, verbose: bool = False, fold: bool = False) -> dict:
    """Time the candidates for a lowered circuit and cache the fastest.

    ``jit`` selects plan-specialised kernels for the timing (default: for
    plans of JQ_JIT_MIN_WIRES or more wires, as the library does)."""
    from jaqsi_amd import lowering

    mats = lowering.matrix_tensor(low.matrix_values, low.n_matrix_elems, None)
    mats = mats.detach() if mats is not None else None
    use_jit = (low.n_wires >= 20) if jit is None else jit
    tried: List[dict] = []

    def measure(low_bits: int, flags: int, tile_bits: int = 0) -> float:
        plan = nat.Plan(low.n_wires, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, dtype,
                        tile_bits, 1, low_bits, flags, fold_tail=fold and not density)
        if use_jit or tile_bits == 14:
            plan.jit(nat.JQ_JIT_SYNC)
        ms = _time_step(plan, batch, len(low.angle_values), mats, density, reps, rounds)
        tried.append({"low_bits": low_bits, "tile_bits": tile_bits, "variant_flags": flags, "ms": round(ms, 4),
                      "passes": plan.info()["n_passes"]})
        if verbose:
            print(f"  low_bits={low_bits} tile_bits={tile_bits} flags={flags}: {ms:.3f} ms "
                  f"({tried[-1]['passes']} passes)", flush=True)
        return ms

    def better(ms: float, best: float) -> bool:
        # a candidate must win by more than the run-to-run spread
        return ms < best * (1.0 - _MIN_GAIN)

    cands = _LOW_CANDIDATES[dtype]
    best_low, best_ms = cands[0], measure(cands[0], 0)
    default_ms = best_ms
    for lb in cands[1:]:
        ms = measure(lb, 0)
        if better(ms, best_ms):
            best_low, best_ms = lb, ms
    # 14-bit tiles (complex64): one 1024-thread workgroup per 2^14 amplitudes
    # (plan-specialised kernels only; wide-pass and dense 3-wire blocks keep
    # 13-bit tiles, and plan creation refuses them)
    best_tile = 0
    if dtype == 0 and low.n_wires >= 14:
        try:
            ms = measure(best_low, 0, 14)
        except (NotImplementedError, ValueError, RuntimeError):
            ms = float("inf")
        if better(ms, best_ms):
            best_tile, best_ms = 14, ms
    best_flags = 0
    # (14-bit plans always run the non-persistent, 32-amplitude backward)
    flags = _FLAG_CANDIDATES[dtype] if best_tile != 14 else (nat.JQ_TUNE_BWD16, nat.JQ_TUNE_NO_DMA_FWD)
    for fl in flags:
        ms = measure(best_low, best_flags | fl, best_tile)
        if better(ms, best_ms):
            best_flags, best_ms = best_flags | fl, ms
    rec = {"low_bits": best_low, "tile_bits": best_tile, "variant_flags": best_flags, "ms": round(best_ms, 4),
           "default_ms": round(default_ms, 4), "batch": batch, "jit": use_jit, "tried": tried}
    _store(entry_key(arch(), low.key, dtype, batch, fold and not density), rec)
    _memo.clear()
    return dict(rec, source="tuned")


def config_for(low, dtype: int, batch: int, density: bool = False, jit: Optional[bool] = None,
               fold: bool = False) -> dict:
    """The knobs plan creation uses for this structure (see the module
    docstring for the modes)."""
    if _mode == "off" or not torch.cuda.is_available():
        return dict(DEFAULT)
    fold = bool(fold) and not density
    mk = (low.key, dtype, batch_bucket(batch), density, _mode, fold)
    hit = _memo.get(mk)
    if hit is not None:
        return dict(hit)
    rec = lookup(low.key, dtype, batch, fold)
    if rec is None and _mode == "tune":
        rec = tune(low, dtype, batch, density, jit=jit, verbose=bool(os.environ.get("JAQSI_AMD_AUTOTUNE_VERBOSE")),
                   fold=fold)
    out = dict(DEFAULT) if rec is None else {"low_bits": int(rec["low_bits"]),
                                               "variant_flags": int(rec["variant_flags"]),
                                               "tile_bits": int(rec.get("tile_bits", 0))}
    if len(_memo) > 256:
        _memo.clear()
    _memo[mk] = out
    return dict(out)


# ---------------------------------------------------------------------------
# tuning the measured workloads: python -m jaqsi_amd.autotune
# ---------------------------------------------------------------------------
def _hea_low(n: int, layers: int, features: bool = False):
    from jaqsi_amd import Gates, lowering
    from jaqsi_amd.tape import recording

    with recording() as tape:
        for layer in range(layers):
            if features and layer == 0:
                for q in range(n):
                    Gates.RX(0.0, wires=q)
            for q in range(n):
                Gates.RX(0.0, wires=q)
                Gates.RY(0.0, wires=q)
                Gates.RZ(0.0, wires=q)
            for q in range(n):
                Gates.CX(wires=[q, (q + 1) % n])
    return lowering.lower_pure(tape, n)


def report(low, dtype: int, batch: int, fold: bool = False) -> dict:
    """The knobs plan creation uses for this structure and where they come
    from ("table", "database", "tuned" or "built-in"), for logs and bench
    records."""
    knobs = config_for(low, dtype, batch, fold=fold)
    rec = lookup(low.key, dtype, batch, fold) if _mode != "off" and torch.cuda.is_available() else None
    out = dict(knobs, source=rec.get("source", "database") if rec else "built-in", mode=_mode)
    if rec and "default_ms" in rec:
        out.update(tuned_ms=rec["ms"], built_in_ms=rec["default_ms"])
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="tune the plan knobs of the statevector workloads on this GPU")
    ap.add_argument("--out", default=None, help="write the chosen entries to this JSON file as well")
    ap.add_argument("--workloads", default="hea24,hea14", help="comma-separated: hea24 (24 qubits, 8 layers, "
                    "batch 1), hea14 (14 qubits, 12 layers, batch 256)")
    args = ap.parse_args(argv)
    picked = {}
    for w in args.workloads.split(","):
        # (the 24-qubit gradient is measured through Z observables: its plan
        # folds the final CX ring)
        if w == "hea24":
            low, batch, fold = _hea_low(24, 8), 1, True
        elif w == "hea14":
            low, batch, fold = _hea_low(14, 12, features=True), 256, False
        else:
            raise SystemExit(f"unknown workload {w!r}")
        print(f"tuning {w}", flush=True)
        rec = tune(low, 0, batch, jit=True, verbose=True, fold=fold)
        picked[entry_key(arch(), low.key, 0, batch, fold)] = {k: v for k, v in rec.items() if k != "source"}
        print(f"{w}: low_bits={rec['low_bits']} variant_flags={rec['variant_flags']} {rec['ms']} ms "
              f"(built-in {rec['default_ms']} ms)", flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(picked, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
