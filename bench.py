"""Benchmark of the adjoint-gradient hot path on one MI355X.

Workload (default): 24-qubit statevector, complex64, hardware-efficient
ansatz with 8 layers (RX, RY, RZ on every wire, then the CX ring
q -> q+1 mod n), 576 angles uniform[0, 2pi) from numpy default_rng(1002);
one step = one forward pass plus the full adjoint gradient of
sum_q <Z_q> w.r.t. all 576 angles, with the angles already resident in HBM.

Prints ONE JSON line with the metric, the roofline of the dominant kernel
(HIP-event timed inside the timed region) and the CPU baseline (the
C++/OpenMP restatement of the reference algorithm timed on this host).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "adjoint-grad circuits/s at 24q & batch 4096×14q; HBM GB/s vs peak"

# Pass kernels: plan-specialised (compiled once per plan before the warm-up,
# outside the timed region) unless --interpreter selects the program
# interpreter kernels.
USE_JIT = True


def kernel_names():
    """Labels of the forward / backward pass kernels (as rocprofv3 names them)."""
    if USE_JIT:
        return ("jq_sv_forward_s*", "jq_sv_backward_s*")
    return ("sv_forward_kernel", "sv_backward_kernel")


def specialise(plan):
    """Compile the plan's specialised pass kernels (when enabled); seconds."""
    from jaqsi_amd import _native as nat

    if not USE_JIT:
        plan.jit(nat.JQ_JIT_OFF)
        return 0.0
    t0 = time.perf_counter()
    plan.jit(nat.JQ_JIT_SYNC)
    return round(time.perf_counter() - t0, 2)
HBM_PEAK_GBS = 8000.0


def median_runs(fn, runs=3, warmup=1):
    """The CPU-baseline protocol: `warmup` untimed runs, then the median
    wall time of `runs` timed ones (seconds) and the individual times."""
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), [round(t, 4) for t in ts]


def host_cpu():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def record_hea(theta, n, features=None):
    from jaqsi_amd import Gates

    if features is not None:
        for q in range(n):
            Gates.RX(float(features[q]), wires=q)
    for layer in range(theta.shape[0]):
        for q in range(n):
            Gates.RX(float(theta[layer, q, 0]), wires=q)
            Gates.RY(float(theta[layer, q, 1]), wires=q)
            Gates.RZ(float(theta[layer, q, 2]), wires=q)
        for q in range(n):
            Gates.CX(wires=[q, (q + 1) % n])


def build_plan(n, layers, features=False, tile_bits=0, fold=True):
    """Plan of the ansatz, as the Script path builds it for Z expectation
    values: the final CX ring is measured through the observables
    (fold_tail), unless fold is False."""
    from jaqsi_amd import autotune, lowering
    from jaqsi_amd import _native as nat
    from jaqsi_amd.tape import recording

    theta = np.zeros((layers, n, 3))
    with recording() as tape:
        record_hea(theta, n, np.zeros(n) if features else None)
    low = lowering.lower_pure(tape, n)
    knobs = autotune.config_for(low, 0, 1, jit=USE_JIT, fold=fold)
    plan = nat.Plan(n, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, 0,
                    tile_bits or knobs.get("tile_bits", 0), 1, knobs["low_bits"], knobs["variant_flags"],
                    fold_tail=fold)
    mats = lowering.matrix_tensor(low.matrix_values, low.n_matrix_elems, None)
    plan.tuning = autotune.report(low, 0, 1, fold=fold)
    return plan, mats


def bench_c2(steps, warmup, n=24, layers=8, fold=True):
    import torch

    from jaqsi_amd import _native as nat

    L = nat.lib()
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(1002)
    theta = rng.uniform(0, 2 * np.pi, (layers, n, 3))
    plan, mats = build_plan(n, layers, fold=fold)
    compile_s = specialise(plan)
    mats = mats.reshape(1, -1).to(dev) if mats is not None else None
    info = plan.info()
    angles = torch.tensor(theta.reshape(1, -1), dtype=torch.float64, device=dev)
    ct = torch.ones((1, n), dtype=torch.float64, device=dev)
    state = torch.empty((1, 1 << n), dtype=torch.complex64, device=dev)
    ev = torch.empty((1, n), dtype=torch.float64, device=dev)
    ga = torch.empty_like(angles)
    wires = np.arange(n, dtype=np.int32)
    d = np.tile(np.array([1.0, -1.0]), n)
    stream = torch.cuda.current_stream(dev).cuda_stream

    def step():
        nat.check(L.jq_sv_adjoint(plan.handle, 1, nat.ptr(angles), nat.ptr(mats), None, 0, nat.ptr(state), nat.JQ_OBS_SV_DIAG,
                                  n, nat.ptr(wires), nat.ptr(d), nat.ptr(ct), nat.ptr(ev), nat.ptr(ga), None, None,
                                  stream))

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    nat.check(L.jq_timing_enable(1))
    for w in range(3):
        nat.check(L.jq_timing_read(w, None, None))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    nat.check(L.jq_timing_enable(0))
    import ctypes as C

    kt = {}
    for w, name in enumerate(kernel_names()):
        tot, cnt = C.c_double(), C.c_int32()
        nat.check(L.jq_timing_read(w, C.byref(tot), C.byref(cnt)))
        kt[name] = (tot.value, cnt.value)
    return {
        "theta": theta,
        "seconds": t1 - t0,
        "info": info,
        "kernels": kt,
        "expvals": ev.cpu().numpy()[0],
        "grads": ga.cpu().numpy()[0],
        "n": n,
        "compile_s": compile_s,
        "tuning": plan.tuning,
        "folded": plan.fold()[0],
        "bwd_fma": bwd_fma([plan]),
    }


def stored_traffic(leg, kernel):
    """HBM bytes per launch of `kernel` from the stored FETCH_SIZE /
    WRITE_SIZE calibration (profiles/traffic.json; its commands are in
    profiles/traffic_*.txt): a measurement of the committed kernels, not of
    this run, so it is labelled as stored."""
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(tf):
        return None
    with open(tf) as f:
        rec = json.load(f).get(leg, {}).get(kernel)
    return rec["total"] if rec else None


def bwd_fma(plans):
    """Backward FMA per amplitude and fused 2x2 block, averaged over the
    blocks of `plans`: 24 (psi <- U^-1 psi, lambda <- U^T lambda, the 4
    complex cotangent entries), 22 for Pauli-rotation blocks (block table
    word 10), whose cotangent the specialised backward keeps as 3 reals."""
    quat = total = 0
    for p in plans:
        bt = p.table(0).reshape(-1, 12)
        total += len(bt)
        quat += int((bt[:, 10] != 0).sum())
    return (22 * quat + 24 * (total - quat)) / max(total, 1)


def roofline_c2(res):
    n = res["n"]
    sweep = 2 * 8 * (1 << n)  # read + write of one complex64 vector
    bytes_per = dict(zip(kernel_names(), (sweep, sweep * 3 // 2)))
    dom = max(res["kernels"], key=lambda k: res["kernels"][k][0])
    tot, cnt = res["kernels"][dom]
    avg_s = tot / 1e3 / max(cnt, 1)
    achieved = bytes_per[dom] / avg_s / 1e9
    traffic = stored_traffic("c2", dom)
    others = {k: {"avg_us": round(v[0] / max(v[1], 1) * 1e3, 2), "launches": v[1]} for k, v in res["kernels"].items()}
    # The same launches against the FP32 VALU: per amplitude and fused 2x2
    # block 8 FMA forward, 22 / 24 backward (bwd_fma), over the plan's blocks per pass
    fma = dict(zip(kernel_names(), (8, res["bwd_fma"])))
    blocks_per_pass = res["info"]["n_blocks"] / res["info"]["n_passes"]
    valu = {}
    for k, (t, c) in res["kernels"].items():
        flop = 2 * fma[k] * blocks_per_pass * (1 << n)
        tf = flop / (t / 1e3 / max(c, 1)) / 1e12
        valu[k] = {"flop_per_launch": round(flop), "achieved_tflops": round(tf, 2),
                   "frac": round(tf / VALU_F32_PEAK_TFLOPS, 4)}
    return {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "algorithmic_bytes_per_launch": bytes_per[dom],
        "kernels": others,
        "valu_fp32": dict(valu, peak_tflops=VALU_F32_PEAK_TFLOPS,
                          note="fused-block FMAs (8 forward / 22 backward per amplitude and Pauli-rotation block, "
                          "averaged over the plan's passes) over the same launch times, against FP32 VALU issue "
                          "without packed instructions; the backward's FMA floor lies above its HBM floor"),
    }


def bench_c3(steps, warmup, B=4096, n=14, seed=1003):
    """Ragged batch: B circuits on n qubits, depth d_b ~ U{2..12}, an RX(x)
    encoding layer with x ~ U[-pi, pi), angles ~ U[0, 2pi); one step = the
    14 Z expvals and the adjoint gradient of sum_q <Z_q> of every circuit."""
    import ctypes as C

    import torch

    from jaqsi_amd import _native as nat
    from jaqsi_amd.ragged import RaggedBatch, record

    L = nat.lib()
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(seed)
    depths = rng.integers(2, 13, size=B)
    feats = rng.uniform(-np.pi, np.pi, (B, n))
    thetas = [rng.uniform(0, 2 * np.pi, (int(d), n, 3)) for d in depths]
    uniq = sorted(set(int(d) for d in depths))
    members = {d: [b for b in range(B) if depths[b] == d] for d in uniq}
    tapes = [record(record_hea, np.zeros((d, n, 3)), n, np.zeros(n)) for d in uniq]
    # one set of tuning knobs for all groups: the deepest group's choice
    from jaqsi_amd import autotune, lowering

    rep = lowering.lower_pure(tapes[-1], n)
    knobs = autotune.config_for(rep, 0, len(members[uniq[-1]]), jit=USE_JIT, fold=True)
    tuning = autotune.report(rep, 0, len(members[uniq[-1]]), fold=True)
    rb = RaggedBatch(tapes, [len(members[d]) for d in uniq], n, [(q, 1.0, -1.0) for q in range(n)], dtype=0,
                     knobs=knobs)
    t_c = time.perf_counter()
    rb.specialise(nat.JQ_JIT_SYNC if USE_JIT else nat.JQ_JIT_OFF)
    compile_s = round(time.perf_counter() - t_c, 2) if USE_JIT else 0.0
    angles = [
        torch.tensor(np.stack([np.concatenate([feats[b], thetas[b].reshape(-1)]) for b in members[d]]), device=dev)
        for d in uniq
    ]
    cts = [torch.ones((len(members[d]), n), dtype=torch.float64, device=dev) for d in uniq]
    for _ in range(warmup):
        rb.adjoint(angles, cts)
    torch.cuda.synchronize()
    nat.check(L.jq_timing_enable(1))
    for w in range(3):
        nat.check(L.jq_timing_read(w, None, None))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        evs, grads = rb.adjoint(angles, cts)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    nat.check(L.jq_timing_enable(0))
    kt = {}
    for w, name in enumerate(kernel_names()):
        tot, cnt = C.c_double(), C.c_int32()
        nat.check(L.jq_timing_read(w, C.byref(tot), C.byref(cnt)))
        kt[name] = (tot.value, cnt.value)
    # The same call recorded once into a HIP graph and replayed: the launch
    # sequence of all groups (side streams joined by events) without host
    # launch overhead.  The replays write the same outputs as the eager call.
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        rb.adjoint(angles, cts)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):  # the stream whose workspaces the warm-up sized
        gevs, ggrads = rb.adjoint(angles, cts)
    graph.replay()
    torch.cuda.synchronize()
    g0 = time.perf_counter()
    for _ in range(steps):
        graph.replay()
    torch.cuda.synchronize()
    g1 = time.perf_counter()
    diff = max(max(float((a - b).abs().max()) for a, b in zip(gevs, evs)),
               max(float((a - b).abs().max()) for a, b in zip(ggrads, grads)))
    circuit_passes = sum(len(members[d]) * p.info()["n_passes"] for d, p in zip(uniq, rb.plans))
    circuit_blocks = sum(len(members[d]) * p.info()["n_blocks"] for d, p in zip(uniq, rb.plans))
    return {
        "compile_s": compile_s,
        "seconds": g1 - g0,
        "eager_seconds": t1 - t0,
        "graph_vs_eager_max_abs_diff": diff,
        "kernels": kt,
        "circuit_passes": circuit_passes,
        "circuit_blocks": circuit_blocks,
        "folded_tail_gates": sorted(set(p.fold()[0] for p in rb.plans)),
        "bwd_fma": bwd_fma(rb.plans),
        "plans": rb.plans,
        "groups": len(uniq),
        "tuning": tuning,
        "depths": depths,
        "feats": feats,
        "thetas": thetas,
        "n": n,
        "B": B,
        "check": (members, uniq, evs, grads),
    }


def roofline_c3(res, steps):
    """C3 is bound by FP32 VALU issue, not HBM: with 14-bit tiles a 14-qubit
    circuit is one tile and most circuits take one or two passes.  `achieved`
    is the algorithmic FP32 FLOPs of the fused 2x2 blocks (per amplitude and
    block 8 FMA forward -- U psi -- and 24 backward -- U^-1 psi, the cotangent
    lambda psi, U^T lambda -- at 2 FLOP per FMA) of every circuit of the timed
    steps over their wall time (graph replays), against the FP32 VALU peak
    without packed instructions (the build keeps packed FP32 off: one FMA per
    lane and cycle, 78.6 TFLOP/s; 157.3 with packed issue).  The HBM figure
    (algorithmic bytes of every pass launch over the same wall time) is kept
    beside it; launches of concurrent groups overlap, so the per-kernel
    entries are summed event times of the eager steps."""
    n = res["n"]
    flops = 2 * (8 + res["bwd_fma"]) * (1 << n) * res["circuit_blocks"] * steps
    ach = flops / res["seconds"] / 1e12
    sweep = 2 * 8 * (1 << n)
    per = dict(zip(kernel_names(), (sweep, sweep * 3 // 2)))
    out = {}
    total_bytes = 0
    for k, (tot, cnt) in res["kernels"].items():
        total_bytes += per[k] * res["circuit_passes"] * steps
        out[k] = {"ms_total_summed": round(tot, 3), "launches": cnt}
    hbm = total_bytes / res["seconds"] / 1e9
    return {
        "bound": "valu-fp32",
        "kernel": "all pass kernels (" + ", ".join(kernel_names()) + ")",
        "achieved": round(ach, 2),
        "peak": VALU_F32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(ach / VALU_F32_PEAK_TFLOPS, 4),
        "peak_packed_fp32": 157.3,
        "traffic": None,
        "flop_note": "per circuit: fused 2x2 blocks x 2^14 amplitudes x (8 + 22) FMA x 2 FLOP (22: the backward "
        "of a Pauli-rotation block), summed over the 4096 circuits of each timed step",
        "hbm": {"achieved": round(hbm, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(hbm / HBM_PEAK_GBS, 4),
                "traffic_per_launch": {k: stored_traffic("c3", k) for k in kernel_names()},
                "bytes_note": "per circuit pass 2*8*2^14 B forward, 3*8*2^14 B backward"},
        "kernels": out,
    }


def cpu_baseline_c3(res, budget_s=15.0):
    from oracle import cpu

    from concurrent.futures import ThreadPoolExecutor

    cores = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    done = 0
    # one circuit per core at a time (the C call releases the GIL); chunks of
    # `cores` circuits until the time budget is used
    with ThreadPoolExecutor(cores) as ex:
        while time.perf_counter() - t0 < budget_s and done < res["B"]:
            hi = min(res["B"], done + cores)
            list(ex.map(lambda b: cpu.hea_grad(res["thetas"][b], res["feats"][b], 1), range(done, hi)))
            done = hi
    dt = time.perf_counter() - t0
    # 1 thread: the first 32 circuits, one after the other (1 warm-up circuit)
    one, _ = median_runs(lambda: [cpu.hea_grad(res["thetas"][b], res["feats"][b], 1) for b in range(32)], 1, 0)
    return {
        "value": round(done / dt, 2),
        "unit": "circuits/s",
        "cores": cores,
        "cpu": host_cpu(),
        "threads_1": {"value": round(32 / one, 2), "unit": "circuits/s", "sample": "circuits 0..31, sequential"},
        "kind": "port",
        "sample": f"the first {done} circuits of the batch, {cores} at a time (one per core; forward + adjoint "
        "each), C++ restatement of the reference plan",
        "seconds": round(dt, 2),
    }


def record_noisy_hea(theta, n, p_dep, gamma):
    """After every gate, on each touched wire in order: DepolarizingChannel
    then AmplitudeDamping."""
    from jaqsi_amd import Gates, noise

    def after(wires):
        for w in wires:
            noise.DepolarizingChannel(p_dep, wires=w)
            noise.AmplitudeDamping(gamma, wires=w)

    for layer in range(theta.shape[0]):
        for q in range(n):
            Gates.RX(float(theta[layer, q, 0]), wires=q)
            after([q])
            Gates.RY(float(theta[layer, q, 1]), wires=q)
            after([q])
            Gates.RZ(float(theta[layer, q, 2]), wires=q)
            after([q])
        for q in range(n):
            Gates.CX(wires=[q, (q + 1) % n])
            after([q, (q + 1) % n])


def bench_c4(steps, warmup, n=12, layers=6, seed=1004):
    """Noisy density matrix: forward + checkpointed adjoint of Tr(Z0 rho)."""
    import ctypes as C

    import torch

    from jaqsi_amd import _native as nat
    from jaqsi_amd import simulation as S
    from jaqsi_amd.tape import recording

    L = nat.lib()
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0, 2 * np.pi, (layers, n, 3))
    with recording() as tape:
        record_noisy_hea(theta, n, 0.01, 0.005)
    low, plan, angles, mats, psi0, psi0_b = S._prepare(tape, n, True, None, None, 0)
    from jaqsi_amd import autotune

    tuning = autotune.report(low, 0, 1)
    compile_s = specialise(plan)
    runner = S._Runner(plan, 1, True, [(0, 1.0, -1.0)], n, False, 0)
    ct = torch.ones((1, 1), dtype=torch.float64, device=angles.device)
    for _ in range(warmup):
        runner.adjoint(angles, mats, None, ct, False)
    torch.cuda.synchronize()
    nat.check(L.jq_timing_enable(1))
    for w in range(3):
        nat.check(L.jq_timing_read(w, None, None))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ev, ga, _, _ = runner.adjoint(angles, mats, None, ct, False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    nat.check(L.jq_timing_enable(0))
    kt = {}
    for w, name in enumerate(kernel_names()):
        tot, cnt = C.c_double(), C.c_int32()
        nat.check(L.jq_timing_read(w, C.byref(tot), C.byref(cnt)))
        kt[name] = (tot.value, cnt.value)
    info = plan.info()
    sweep = 2 * 8 * (1 << (2 * n))
    per = dict(zip(kernel_names(), (sweep, sweep * 3 // 2)))
    kern = {k: {"avg_us": round(v[0] / max(v[1], 1) * 1e3, 2), "launches": v[1],
                "GB/s": round(per[k] * v[1] / max(v[0] / 1e3, 1e-12) / 1e9, 1)} for k, v in kt.items()}
    dom = max(kt, key=lambda k: kt[k][0])
    return {
        "metric": "noisy density-matrix adjoint-grad evaluations/s (12 qubits, 6 layers, 720 Kraus channels)",
        "value": round(steps / (t1 - t0), 3),
        "unit": "gradient evaluations/s",
        "steps": steps,
        "ms_per_step": round((t1 - t0) / steps * 1e3, 3),
        "dtype": "c64",
        "config": {
            "workload": "rho 4096x4096 complex64 as a 24-wire vector; HEA 6 layers (216 angles U[0,2pi), seed 1004), "
            "DepolarizingChannel(0.01) + AmplitudeDamping(0.005) after every gate on each touched wire; "
            "Tr(Z0 rho) and its gradient by the checkpointed adjoint",
            "passes": info["n_passes"],
            "blocks": info["n_blocks"],
            "pass_kernels": "plan-specialised" if USE_JIT else "interpreter",
            "specialise_compile_s": compile_s,
            "autotune": tuning,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": kern[dom]["GB/s"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(kern[dom]["GB/s"] / HBM_PEAK_GBS, 4),
            "traffic": stored_traffic("c4", dom),
            "algorithmic_bytes_per_launch": per[dom],
            "kernels": kern,
        },
        "expval": float(ev.cpu()[0, 0]),
    }


FP64_PEAK_TFLOPS = 78.6
# FP32 vector issue without packed instructions (one FMA per lane and cycle):
# the build keeps packed FP32 off, so this is the VALU ceiling of the pass kernels
VALU_F32_PEAK_TFLOPS = 78.6
# 64 x (5.892e7 ADD + 1.079e8 MUL + 2 x 1.661e8 FMA) F64 VALU instructions per launch at 8192 sets x 2000 steps
C5_EXECUTED_FP64_FLOP = 64 * (5.892e7 + 1.079e8 + 2 * 1.661e8)


def c5_inputs(n_sets, seed):
    """Parameter sets (A, sigma, beta, phi), frequencies (w0, w1) and the
    target RX(pi/2) x I, drawn with numpy default_rng(seed)."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(0.1, 1.0, n_sets)
    sigma = rng.uniform(5.0, 20.0, n_sets)
    beta = rng.uniform(-1.0, 1.0, n_sets)
    phi = rng.uniform(0.0, 2 * np.pi, n_sets)
    omega = rng.uniform(4.5, 5.5, (n_sets, 2)) * 2 * np.pi
    c, s_ = np.cos(np.pi / 4), np.sin(np.pi / 4)
    rx = np.array([[c, -1j * s_], [-1j * s_, c]])
    return np.stack([A, sigma, beta, phi], axis=1), omega, np.kron(rx, np.eye(2))


def bench_c5(steps, warmup, n_sets=8192, n_steps=2000, seed=1005):
    """Batched pulse propagators: 8192 DRAG-Gaussian parameter sets, 2 coupled
    qubits, 2000 Magnus-2 steps over 50 ns in FP64; infidelity against
    RX(pi/2) x I and its gradient w.r.t. (A, sigma, beta, phi)."""
    import torch

    from jaqsi_amd import _native as nat
    params, omega, V = c5_inputs(n_sets, seed)
    dev = torch.device("cuda", 0)
    P = torch.tensor(np.concatenate([params, omega], axis=1), dtype=torch.float64, device=dev).contiguous()
    Vh = np.ascontiguousarray(np.stack([V.real, V.imag], axis=-1).reshape(-1))
    U = torch.empty((n_sets, 4, 4), dtype=torch.complex128, device=dev)
    inf = torch.empty(n_sets, dtype=torch.float64, device=dev)
    g = torch.empty((n_sets, 4), dtype=torch.float64, device=dev)
    d = nat.PulseDesc(n_sets, n_steps, 50.0, 2 * np.pi * 0.005, nat.JQ_ENV_DRAG_GAUSSIAN, 0)
    stream = torch.cuda.current_stream(dev)

    def step():
        nat.check(nat.lib().jq_pulse_propagate(d, nat.ptr(P), nat.ptr(Vh), nat.ptr(U), nat.ptr(inf), nat.ptr(g),
                                               stream.cuda_stream))

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kms = e0.elapsed_time(e1) / steps
    flops = C5_EXECUTED_FP64_FLOP * (n_sets / 8192) * (n_steps / 2000)
    ach = flops / (kms / 1e3) / 1e12
    return {
        "metric": "pulse-propagator gradient evaluations/s (8192 sets, 2 qubits, 2000 Magnus-2 steps, FP64)",
        "value": round(n_sets * steps / (t1 - t0), 1),
        "unit": "parameter sets/s (propagator + infidelity + 4-parameter gradient)",
        "steps": steps,
        "ms_per_step": round((t1 - t0) / steps * 1e3, 3),
        "dtype": "f64",
        "config": {
            "workload": "8192 sets: A~U[0.1,1], sigma~U[5,20] ns, beta~U[-1,1], phi~U[0,2pi), w_q~U[4.5,5.5]*2pi "
            "rad/ns (seed 1005); J=2pi*0.005 rad/ns; T=50 ns, 2000 steps; H(t)=sum_q w_q/2 Z_q + J Z0Z1 + "
            "DRAG drive on qubit 0; infidelity vs RX(pi/2) x I",
        },
        "roofline": {
            "bound": "fp64",
            "kernel": "pulse_kernel",
            "achieved": round(ach, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(ach / FP64_PEAK_TFLOPS, 4),
            "traffic": None,
            "flop_count": "executed FP64 FLOP per launch: 64 lanes x (ADD + MUL + 2 FMA) F64 VALU instructions "
            "counted by SQ_INSTS_VALU_{ADD,MUL,FMA}_F64 at 8192 sets x 2000 steps (profiles/counters_c5_pulse_kernel.txt), "
            "scaled linearly to this launch; F64 transcendentals (SQ_INSTS_VALU_TRANS_F64) are not counted",
            "kernel_ms": round(kms, 4),
        },
        "sample_infidelity": float(inf[0].cpu()),
    }


def cpu_baseline_c5(n_sample=2048):
    """The C++/OpenMP port of the reference's Magnus-2 solve (one 4x4 expm per
    step) with the infidelity and its adjoint gradient w.r.t. (A, sigma,
    beta, phi): the same work as the GPU leg, on a bounded sample of sets."""
    from oracle import cpu, pulse

    params, omega = pulse.workload(8192, 1005)
    V = pulse.target_rx_half_pi()
    cores = len(os.sched_getaffinity(0))
    P, W = params[:n_sample], omega[:n_sample]
    dt, ts = median_runs(lambda: cpu.pulse_grad(P, W, pulse.N_C5, pulse.T_C5, pulse.J_C5, V, cores), 3, 1)
    one, _ = median_runs(lambda: cpu.pulse_grad(P[:64], W[:64], pulse.N_C5, pulse.T_C5, pulse.J_C5, V, 1), 1, 0)
    return {
        "value": round(n_sample / dt, 2),
        "unit": "parameter sets/s (propagator + infidelity + 4-parameter gradient)",
        "cores": cores,
        "cpu": host_cpu(),
        "kind": "port",
        "sample": f"the first {n_sample} of the 8192 sets, 2000 Magnus-2 steps each (FP64 expm per step, adjoint "
        "gradient through the step exponentials' Frechet derivatives); 1 warm-up, median of 3",
        "seconds": round(dt, 3),
        "runs_s": ts,
        "threads_1": {"value": round(64 / one, 2), "unit": "parameter sets/s", "sample": "sets 0..63, 1 thread"},
    }


def cpu_baseline_c4(theta):
    """The C++/OpenMP port of the reference's density-matrix plan (fused
    superoperator sweeps, CX pairs as gathers), FORWARD ONLY: the reference
    differentiates it by autodiff over taped rho's, which the port does not
    restate."""
    from oracle import cpu

    cores = len(os.sched_getaffinity(0))
    dt, ts = median_runs(lambda: cpu.dm_forward(theta, 0.01, 0.005, cores), 3, 1)
    return {
        "value": round(1.0 / dt, 4),
        "unit": "forward evaluations/s (no gradient)",
        "cores": cores,
        "cpu": host_cpu(),
        "kind": "port",
        "sample": "one forward of the full 12-qubit noisy workload (rho 4096^2 complex64, 288 superoperator sweeps); "
        "1 warm-up, median of 3; forward only",
        "seconds": round(dt, 3),
        "runs_s": ts,
    }


def bench_c1(steps=50, warmup=5, n=10, layers=5, seed=1001):
    """The reference's CPU-runnable case: 10 qubits, 5 layers, 150 angles,
    10 Z expvals, batch 1; latency of one forward + adjoint gradient (ct =
    ones), through Script.execute + backward and through the C-ABI."""
    import torch

    import jaqsi_amd as J
    from jaqsi_amd import _native as nat

    rng = np.random.default_rng(seed)
    theta = rng.uniform(0, 2 * np.pi, (layers, n, 3))
    J.enable_x64(False)
    th = torch.tensor(theta, dtype=torch.float64, device="cuda", requires_grad=True)
    obs = [J.PauliZ(wires=q, record=False) for q in range(n)]
    script = J.Script(lambda t: record_hea_t(t, n), n_qubits=n)

    def script_step():
        th.grad = None
        ev = script.execute(type="expval", obs=obs, args=(th,))
        ev.sum().backward()
        return ev

    for _ in range(warmup):
        script_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        script_step()
    torch.cuda.synchronize()
    script_ms = (time.perf_counter() - t0) / steps * 1e3
    plan, mats = build_plan(n, layers)
    plan.jit(nat.JQ_JIT_OFF)
    dev = torch.device("cuda", 0)
    mats = mats.reshape(1, -1).to(dev) if mats is not None else None
    angles = torch.tensor(theta.reshape(1, -1), dtype=torch.float64, device=dev)
    ct = torch.ones((1, n), dtype=torch.float64, device=dev)
    state = torch.empty((1, 1 << n), dtype=torch.complex64, device=dev)
    ev = torch.empty((1, n), dtype=torch.float64, device=dev)
    ga = torch.empty_like(angles)
    wires = np.arange(n, dtype=np.int32)
    d = np.tile(np.array([1.0, -1.0]), n)
    stream = torch.cuda.current_stream(dev).cuda_stream

    def capi_step():
        nat.check(nat.lib().jq_sv_adjoint(plan.handle, 1, nat.ptr(angles), nat.ptr(mats), None, 0, nat.ptr(state),
                                          nat.JQ_OBS_SV_DIAG, n, nat.ptr(wires), nat.ptr(d), nat.ptr(ct), nat.ptr(ev),
                                          nat.ptr(ga), None, None, stream))

    for _ in range(warmup):
        capi_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        capi_step()
    torch.cuda.synchronize()
    capi_ms = (time.perf_counter() - t0) / steps * 1e3
    out = {
        "metric": "forward + adjoint latency (10 qubits, 5 layers, 150 angles, batch 1)",
        "value": round(script_ms, 4),
        "unit": "ms per gradient (Script.execute + backward)",
        "higher_is_better": False,
        "capi_ms": round(capi_ms, 4),
        "steps": steps,
        "dtype": "c64",
        "config": {"workload": "statevector 10 qubits, HEA 5 layers (RX,RY,RZ per wire + CX ring), 150 angles "
                   "uniform[0,2pi) seed 1001, 10 Z expvals, gradient of their sum; batch 1"},
    }
    from oracle import cpu

    cores = len(os.sched_getaffinity(0))
    one, ts1 = median_runs(lambda: cpu.hea_grad(theta, None, 1), runs=5, warmup=1)
    allc, tsa = median_runs(lambda: cpu.hea_grad(theta, None, cores), runs=5, warmup=1)
    out["cpu_baseline"] = {
        "value": round(one * 1e3, 4),
        "unit": "ms per gradient, 1 thread",
        "cores": 1,
        "cpu": host_cpu(),
        "kind": "port",
        "sample": "the full C1 workload: C++ restatement of the reference plan, forward + adjoint; 1 warm-up, "
        "median of 5",
        "all_cores": {"value": round(allc * 1e3, 4), "unit": "ms per gradient", "cores": cores, "runs_s": tsa},
        "runs_s": ts1,
    }
    return out


def record_hea_t(t, n):
    from jaqsi_amd import Gates

    for layer in range(t.shape[0]):
        for q in range(n):
            Gates.RX(t[layer, q, 0], wires=q)
            Gates.RY(t[layer, q, 1], wires=q)
            Gates.RZ(t[layer, q, 2], wires=q)
        for q in range(n):
            Gates.CX(wires=[q, (q + 1) % n])


def script_path_c2(theta, steps=20, warmup=2):
    """The headline step as a user runs it: Script.execute(...) on the
    resident angles, then .sum().backward() (the backward consumes the
    checkpoints the forward left in the plan).  Steady state of a training
    loop: the GPU time between the end of a lead step and the end of the last
    of `steps` steps (events on the launch stream), so the host recording of
    the first step -- during which the GPU has nothing queued yet -- is not
    charged to every step; host time that does not keep ahead of the GPU
    shows up as idle time between the events."""
    import torch

    import jaqsi_amd as J
    from jaqsi_amd import _native as nat
    from jaqsi_amd import simulation

    n = theta.shape[1]
    J.enable_x64(False)
    simulation.specialise(nat.JQ_JIT_SYNC if USE_JIT else nat.JQ_JIT_OFF)
    th = torch.tensor(theta, dtype=torch.float64, device="cuda", requires_grad=True)
    obs = [J.PauliZ(wires=q, record=False) for q in range(n)]
    script = J.Script(lambda t: record_hea_t(t, n), n_qubits=n)

    def step():
        th.grad = None
        script.execute(type="expval", obs=obs, args=(th,)).sum().backward()

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    step()
    e0.record()
    for _ in range(steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    simulation.specialise(None)
    return e0.elapsed_time(e1) / steps


def cpu_baseline_c2(theta):
    from oracle import cpu

    cores = len(os.sched_getaffinity(0))
    dt, ts = median_runs(lambda: cpu.hea_grad(theta, None, cores), runs=3, warmup=1)
    return {
        "value": round(1.0 / dt, 5),
        "unit": "gradient evaluations/s",
        "cores": cores,
        "kind": "port",
        "cpu": host_cpu(),
        "sample": "forward + adjoint gradient of the full 24-qubit, 8-layer workload (complex64), C++/OpenMP "
        "restatement of the reference plan (same-wire fusion, composed CX-ring gather); 1 warm-up, median of 3",
        "seconds": round(dt, 2),
        "runs_s": ts,
        "threads_1": "not timed by default: one single-thread gradient takes minutes at 24 qubits (the C1 and C3 "
        "legs carry 1-thread figures of the same code)",
    }


def naive_hea_expval(theta, n):
    """The plain GPU path the speed-up target is stated against: gate-by-gate
    complex64 statevector in PyTorch (RX, RY, RZ on every wire in the slice
    form of the reference's _apply_gate, then the CX ring), the gradient by
    autograd taping every intermediate state.  Returns sum_q <Z_q>."""
    import torch

    dev = theta.device
    psi = torch.zeros(1 << n, dtype=torch.complex64, device=dev)
    psi[0] = 1
    eye = torch.eye(2, dtype=torch.complex64, device=dev)
    X = torch.tensor([[0, 1], [1, 0]], dtype=torch.complex64, device=dev)
    Y = torch.tensor([[0, -1j], [1j, 0]], dtype=torch.complex64, device=dev)
    Z = torch.tensor([[1, 0], [0, -1]], dtype=torch.complex64, device=dev)

    def rot(P, t):  # cos(t/2) I - i sin(t/2) P, differentiable in t
        c, sn = torch.cos(t / 2).to(torch.complex64), torch.sin(t / 2).to(torch.complex64)
        return c * eye - 1j * sn * P

    def apply1(psi, G, q):  # wire q = bit n-1-q; out_i = sum_j G[i, j] slice_j (simulation.py:70-112)
        v = psi.reshape(1 << q, 2, 1 << (n - 1 - q))
        v0, v1 = v[:, 0, :], v[:, 1, :]
        return torch.stack([G[0, 0] * v0 + G[0, 1] * v1, G[1, 0] * v0 + G[1, 1] * v1], dim=1).reshape(-1)

    def cx(psi, c, t):
        v = psi.reshape([2] * n)
        flipped = torch.flip(v, dims=[t])
        idx = [slice(None)] * n
        idx[c] = 1
        out = v.clone()
        out[tuple(idx)] = flipped[tuple(idx)]
        return out.reshape(-1)

    for layer in range(theta.shape[0]):
        for q in range(n):
            psi = apply1(psi, rot(X, theta[layer, q, 0]), q)
            psi = apply1(psi, rot(Y, theta[layer, q, 1]), q)
            psi = apply1(psi, rot(Z, theta[layer, q, 2]), q)
        for q in range(n):
            psi = cx(psi, q, (q + 1) % n)
    p = (psi.real ** 2 + psi.imag ** 2).reshape([2] * n)
    tot = 0
    for q in range(n):
        m = p.sum(dim=[d for d in range(n) if d != q])
        tot = tot + (m[0] - m[1])
    return tot


def naive_hea_batched(theta, n, features=None):
    """The same plain path for a batch of circuits of one structure (the
    reference vmaps a structure group): state (B, 2^n), the RX(x) encoding
    layer when features (B, n) are given, then the ansatz layers of theta
    (B, layers, n, 3).  Returns sum_q <Z_q> per circuit, shape (B,)."""
    import torch

    dev = theta.device
    B = theta.shape[0]
    psi = torch.zeros((B, 1 << n), dtype=torch.complex64, device=dev)
    psi[:, 0] = 1

    def rot(axis, t):  # (B, 2, 2) rotations cos(t/2) I - i sin(t/2) P
        c, sn = torch.cos(t / 2).to(torch.complex64), torch.sin(t / 2).to(torch.complex64)
        z = torch.zeros_like(c)
        if axis == 0:
            rows = [[c, -1j * sn], [-1j * sn, c]]
        elif axis == 1:
            rows = [[c, -sn], [sn, c]]
        else:
            rows = [[c - 1j * sn, z], [z, c + 1j * sn]]
        return rows

    def apply1(psi, G, q):
        v = psi.reshape(B, 1 << q, 2, 1 << (n - 1 - q))
        v0, v1 = v[:, :, 0, :], v[:, :, 1, :]
        g = [[x[:, None, None] for x in r] for r in G]
        return torch.stack([g[0][0] * v0 + g[0][1] * v1, g[1][0] * v0 + g[1][1] * v1], dim=2).reshape(B, -1)

    def cx(psi, c, t):
        v = psi.reshape([B] + [2] * n)
        flipped = torch.flip(v, dims=[1 + t])
        idx = [slice(None)] * (n + 1)
        idx[1 + c] = 1
        out = v.clone()
        out[tuple(idx)] = flipped[tuple(idx)]
        return out.reshape(B, -1)

    if features is not None:
        for q in range(n):
            psi = apply1(psi, rot(0, features[:, q]), q)
    for layer in range(theta.shape[1]):
        for q in range(n):
            for axis in range(3):
                psi = apply1(psi, rot(axis, theta[:, layer, q, axis]), q)
        for q in range(n):
            psi = cx(psi, q, (q + 1) % n)
    p = (psi.real ** 2 + psi.imag ** 2).reshape([B] + [2] * n)
    tot = 0
    for q in range(n):
        m = p.sum(dim=[1 + d for d in range(n) if d != q])
        tot = tot + (m[:, 0] - m[:, 1])
    return tot


def naive_gpu_c3(res, steps=2):
    """The plain path over the ragged batch: one batched autograd program per
    depth group, groups one after the other."""
    import torch

    n, depths, feats, thetas = res["n"], res["depths"], res["feats"], res["thetas"]
    groups = []
    for d in sorted(set(int(x) for x in depths)):
        idx = [b for b in range(len(depths)) if depths[b] == d]
        th = torch.tensor(np.stack([thetas[b] for b in idx]), dtype=torch.float32, device="cuda", requires_grad=True)
        x = torch.tensor(feats[idx], dtype=torch.float32, device="cuda")
        groups.append((th, x))

    def step():
        for th, x in groups:
            th.grad = None
            naive_hea_batched(th, n, x).sum().backward()

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return {
        "value": round(len(depths) / dt, 1),
        "unit": "circuits/s",
        "kind": "plain PyTorch on the same GPU",
        "sample": f"the full 4096-circuit batch, one batched autograd program per depth group (gates in the "
        f"reference's slice form, CX as flips), complex64; {steps} timed batch steps after one warm-up",
        "seconds_per_batch": round(dt, 4),
    }


def naive_gpu_c2(theta, steps=2):
    import torch

    n = theta.shape[1]
    th = torch.tensor(theta, dtype=torch.float32, device="cuda", requires_grad=True)
    naive_hea_expval(th, n).backward()  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        th.grad = None
        naive_hea_expval(th, n).backward()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return {
        "value": round(1.0 / dt, 3),
        "unit": "gradient evaluations/s",
        "kind": "plain PyTorch on the same GPU",
        "sample": f"the full 24-qubit, 8-layer workload, complex64: every gate applied to the whole state in the "
        f"reference's slice form (out_i = sum_j G_ij slice_j), the CX ring as flips, autograd over the taped states; "
        f"{steps} timed gradients after one warm-up",
        "seconds_per_gradient": round(dt, 4),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the 4096x14-qubit ragged-batch leg")
    ap.add_argument("--c3-steps", type=int, default=5)
    ap.add_argument("--no-c4", action="store_true", help="skip the 12-qubit noisy density-matrix leg")
    ap.add_argument("--no-c5", action="store_true", help="skip the pulse-propagator leg")
    ap.add_argument("--no-naive", action="store_true", help="skip the plain-PyTorch GPU comparison of the headline")
    ap.add_argument(
        "--legs",
        default="c1,c2,c3,c4,c5",
        help="comma-separated workloads to run (profiling runs select one); the headline line needs c2",
    )
    ap.add_argument("--autotune", default="lookup", choices=["off", "lookup", "tune"],
                    help="plan knobs: built-in (off), the tuned tables (lookup) or tune now on a table miss (tune)")
    ap.add_argument("--interpreter", action="store_true",
                    help="run the program-interpreter pass kernels instead of the plan-specialised ones")
    ap.add_argument(
        "--copy-probe",
        type=int,
        default=0,
        metavar="MIB",
        help="only run a plain device copy of MIB MiB three times (calibrates the HBM byte counters)",
    )
    args = ap.parse_args()
    global USE_JIT
    USE_JIT = not args.interpreter
    from jaqsi_amd import autotune

    autotune.set_mode(args.autotune)
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: every run in this build uses a single MI355X; "
            "the multi-GPU design (batch / qubit sharding over RCCL) is described in DESIGN.md",
            file=sys.stderr,
        )
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_probe:
        src = torch.ones(args.copy_probe << 18, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe << 20}), flush=True)
        return
    legs = set(args.legs.split(","))
    if "c2" not in legs:
        # profiling run of the secondary legs only: one JSON object per leg
        out = {}
        if "c3" in legs:
            r3 = bench_c3(args.c3_steps, 1)
            out["c3"] = {"circuits_per_s": round(r3["B"] * args.c3_steps / r3["seconds"], 1),
                         "roofline": roofline_c3(r3, args.c3_steps)}
        if "c4" in legs:
            out["c4"] = bench_c4(args.steps, 1)
        if "c5" in legs:
            out["c5"] = bench_c5(args.steps, 1)
        print(json.dumps(out), flush=True)
        return
    res = bench_c2(args.steps, args.warmup)
    ms = res["seconds"] / args.steps * 1e3
    value = args.steps / res["seconds"]
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "gradient evaluations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {
            "workload": "statevector 24 qubits, hardware-efficient ansatz 8 layers (RX,RY,RZ per wire + CX ring), "
            "576 angles uniform[0,2pi) seed 1002, forward + full adjoint gradient of sum_q <Z_q>",
            "n_qubits": 24,
            "layers": 8,
            "params": 576,
            "passes": res["info"]["n_passes"],
            "tile_bits": res["info"]["tile_bits"],
            "folded_tail_gates": res["folded"],
            "pass_kernels": "plan-specialised" if USE_JIT else "interpreter",
            "specialise_compile_s": res["compile_s"],
            "autotune": res["tuning"],
        },
        "roofline": roofline_c2(res),
    }
    print(f"c2 leg done: {value:.3f} gradient evaluations/s", file=sys.stderr, flush=True)
    # the same step with every gate in the passes (the final CX ring
    # simulated instead of measured through the observables)
    uf = bench_c2(10, 2, fold=False)
    line["unfolded"] = {
        "value": round(10 / uf["seconds"], 3),
        "ms_per_step": round(uf["seconds"] / 10 * 1e3, 4),
        "passes": uf["info"]["n_passes"],
        "max_abs_diff_expvals": float(np.max(np.abs(uf["expvals"] - res["expvals"]))),
        "max_abs_diff_grads": float(np.max(np.abs(uf["grads"] - res["grads"]))),
        "note": "the plan with fold_tail off: the final CX ring runs in the passes (10 timed steps)",
    }
    sp_ms = script_path_c2(res["theta"])
    line["script_path"] = {
        "ms_per_step": round(sp_ms, 4),
        "vs_capi_step": round(sp_ms / ms, 4),
        "note": "Script.execute(type='expval', ...) then .sum().backward() on the same workload: one forward "
        "(checkpoints kept in the plan) and one backward sweep, plus the host recording and lowering; steady "
        "state over 20 steps (GPU events after a lead step and after the last step)",
    }
    print("c2 script path done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_c2(res["theta"])
        print("c2 cpu baseline done", file=sys.stderr, flush=True)
    if not args.no_naive:
        nb = naive_gpu_c2(res["theta"])
        nb["speedup"] = round(value / nb["value"], 1)
        line["naive_gpu_baseline"] = nb
        print("naive gpu comparison done", file=sys.stderr, flush=True)
    if not args.no_c3 and "c3" in legs:
        r3 = bench_c3(args.c3_steps, 1)
        c3 = {
            "metric": "ragged-batch adjoint-grad circuits/s (4096 x 14 qubits, depth U{2..12})",
            "value": round(r3["B"] * args.c3_steps / r3["seconds"], 1),
            "unit": "circuits/s",
            "steps": args.c3_steps,
            "ms_per_step": round(r3["seconds"] / args.c3_steps * 1e3, 3),
            "dtype": "c64",
            "config": {
                "workload": "4096 circuits, 14 qubits, RX(x) encoding + d_b in U{2..12} ansatz layers "
                "(RX,RY,RZ per wire + CX ring), x ~ U[-pi,pi), angles ~ U[0,2pi), seed 1003; "
                "14 Z expvals + adjoint gradient of sum_q <Z_q> per circuit",
                "groups": r3["groups"],
                "autotune": r3["tuning"],
                "circuit_passes": r3["circuit_passes"],
                "folded_tail_gates": r3["folded_tail_gates"],
                "circuit_blocks": r3["circuit_blocks"],
                "tile_bits": sorted({p.info()["tile_bits"] for p in r3["plans"]}),
                "pass_kernels": "plan-specialised (one launch sequence per group)" if USE_JIT
                else "interpreter (one launch per pass step over all groups)",
                "specialise_compile_s": r3["compile_s"],
                "execution": "HIP graph replay of the captured ragged-batch adjoint call",
                "eager_circuits_per_s": round(r3["B"] * args.c3_steps / r3["eager_seconds"], 1),
                "graph_vs_eager_max_abs_diff": r3["graph_vs_eager_max_abs_diff"],
            },
            "roofline": roofline_c3(r3, args.c3_steps),
        }
        if not args.no_cpu:
            c3["cpu_baseline"] = cpu_baseline_c3(r3)
        if not args.no_naive:
            nb = naive_gpu_c3(r3)
            nb["speedup"] = round(c3["value"] / nb["value"], 1)
            c3["naive_gpu_baseline"] = nb
        line["c3"] = c3
        print("c3 leg done", file=sys.stderr, flush=True)
    if not args.no_c4 and "c4" in legs:
        c4 = bench_c4(3, 1)
        if not args.no_cpu:
            th4 = np.random.default_rng(1004).uniform(0, 2 * np.pi, (6, 12, 3))
            c4["cpu_baseline"] = cpu_baseline_c4(th4)
        line["c4"] = c4
        print("c4 leg done", file=sys.stderr, flush=True)
    if "c1" in legs:
        line["c1"] = bench_c1()
        print("c1 leg done", file=sys.stderr, flush=True)
    if not args.no_c5 and "c5" in legs:
        c5 = bench_c5(3, 1)
        if not args.no_cpu:
            c5["cpu_baseline"] = cpu_baseline_c5()
        line["c5"] = c5
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
