"""Benchmark of the FFCM mobility apply on one MI355X.

One step is one device-resident mobility apply (wrap, hash, radix sort, gather, spread, FFT
Stokes solve, interpolation, correction) of the dense suspension N = 2^20 spheres, L = 280,
grid 512^3, phi = 0.20 with half of the particles in 64 Gaussian clusters (std 8; 10,000
rejected cluster draws before a uniform fallback), M_G = 8, Sigma/sigma = 2, R_c = 4, Float32,
with positions and forces already in HBM. Prints one JSON line: particles/s, ms per apply (wall
clock over the timed steps, and min / median of HIP-event-timed applies), per-stage device time
from HIP events recorded on the library stream, the HBM roofline of the dominant stage, the
accuracy against the CPU Float32 and Float64 restatements of the reference on the same inputs,
and that restatement's throughput on the host cores. Secondary legs (same line): the
65,536-particle / 256^3 suspension, the pullback there, the position-cached operator, the
host-staged apply, the batched 64-system ensemble, and the published cuFCM comparison
configuration.

    python bench.py [--steps K] [--warmup W] [--no-cpu] [--no-extra]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ffcm.jl_amd"))
sys.path.insert(0, ROOT)

METRIC = "particles/s per mobility apply (plain vs batched/adjoint) at rel-L2≤1e-5"
HBM_PEAK_GBS = 8000.0
HEADLINE = 2  # the 2^20-particle / 512^3 suspension


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_threads():
    """Host cores this process may use, at most 16 (the GPU box's share per GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def algorithmic_bytes(M, N, C, key_passes):
    """Per-apply algorithmic HBM bytes of each stage (Float32):
    FFT: 3 (8 M^3 + 64 H), five read+write passes over the three fields;
    interpolation: 12 M^3 + 28 N; spread: 12 M^3 + 24 N;
    cell list: 28 N + 16 P N + 8 C + 52 N (P radix passes)."""
    M3 = M ** 3
    H = (M // 2 + 1) * M * M
    return {
        "stokes_fft": 3 * (8 * M3 + 64 * H),
        "interpolate": 12 * M3 + 28 * N,
        "spread": 12 * M3 + 24 * N,
        "cell_list": 28 * N + 16 * key_passes * N + 8 * C + 52 * N,
    }


def sha256(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def event_times(cfg, fn, reps):
    """Per-call device time (ms) of fn(), each bracketed by HIP events on the library stream."""
    import torch

    stream = torch.cuda.ExternalStream(cfg.stream())
    pairs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in pairs:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    return [a.elapsed_time(b) for a, b in pairs]


def run_plain(kw, Y, F, steps, warmup, label, cached=False):
    """Device-resident applies of one suspension: warm-up, a stage-timed pass, HIP-event-timed
    applies, then exactly `steps` applies timed by the wall clock between synchronisations."""
    import torch

    from ForceCouplingMethod import (FFCMConfig, apply_forces_device_, mobility_, mobility_device_,
                                     set_positions_device_)

    N = Y.shape[1]
    cfg = FFCMConfig(gpu_acceleration=True, **kw)
    dY = torch.from_numpy(np.ascontiguousarray(Y.T)).cuda()
    dF = torch.from_numpy(np.ascontiguousarray(F.T)).cuda()
    dV = torch.empty_like(dY)
    apply = lambda: mobility_device_(dV, cfg, dY, dF, synchronize=False)
    torch.cuda.synchronize()
    for _ in range(warmup):
        apply()
    torch.cuda.synchronize()
    cfg.set_timing(True)
    for _ in range(steps):
        apply()
    torch.cuda.synchronize()
    st = cfg.timing()
    cfg.set_timing(False)
    ev = event_times(cfg, apply, max(steps, 20))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        apply()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    V = dV.cpu().numpy().T.copy()
    log(f"{label}: {dt * 1e3:.3f} ms per apply (wall), event median {np.median(ev):.3f} ms")
    out = {"cfg": cfg, "N": N, "s_per_apply": dt, "event_ms": ev, "V": V,
           "stage_ms": {k: v / max(1, st["calls"]) for k, v in st["ms"].items()}, "info": cfg.info(),
           "dev": (dY, dF, dV)}
    # host-staged apply (PCIe-inclusive), never the headline
    Vh = np.empty((3, N), np.float32, order="F")
    mobility_(Vh, cfg, Y, F)
    reps = max(3, min(steps, 10))
    t0 = time.perf_counter()
    for _ in range(reps):
        mobility_(Vh, cfg, Y, F)
    out["s_per_apply_staged"] = (time.perf_counter() - t0) / reps
    out["staged_vs_device_rel_l2"] = rel_l2(Vh, V)
    if cached:
        # the iterative-solver operator: positions cached once, forces applied per mul!
        set_positions_device_(cfg, dY)
        cached_apply = lambda: apply_forces_device_(dV, cfg, dF, synchronize=False)
        for _ in range(max(3, warmup // 4)):
            cached_apply()
        evc = event_times(cfg, cached_apply, max(steps, 20))
        out["cached"] = {"ms_median": float(np.median(evc)), "ms_min": float(np.min(evc)),
                         "rel_l2_vs_uncached": rel_l2(dV.cpu().numpy().T, V)}
        # the same apply replayed from a captured hipGraph (ffcm_set_graph)
        cfg.set_graph(True)
        for _ in range(3):  # launched, captured, replayed
            apply()
        evg = event_times(cfg, apply, max(steps, 20))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            apply()
        torch.cuda.synchronize()
        dtg = (time.perf_counter() - t0) / steps
        out["graph"] = {"ms_per_apply": dtg * 1e3, "ms_median": float(np.median(evg)), "ms_min": float(np.min(evg)),
                        "max_abs_diff_vs_launched": float(np.max(np.abs(dV.cpu().numpy().T - V)))}
        cfg.set_graph(False)
    return out


def stage_table(r, M):
    N = r["N"]
    cfg = r["cfg"]
    C = int(np.prod(cfg.num_cells))
    key_bits = max(1, math.ceil(math.log2(C)))
    ab = algorithmic_bytes(M, N, C, math.ceil(key_bits / 8))
    stages = {}
    for k, ms in r["stage_ms"].items():
        entry = {"ms": round(ms, 4)}
        if k in ab:
            gbs = ab[k] / (ms * 1e-3) / 1e9
            entry.update({"algorithmic_bytes": ab[k], "GB_per_s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)})
        stages[k] = entry
    return stages


def oracle_apply(kw, Y, F, T, fft_threads, particle_threads):
    from oracle import oracle as orc

    def conv(v):
        if isinstance(v, tuple):
            return tuple(T(x) for x in v)
        return T(v) if isinstance(v, (float, np.floating)) else v

    kwT = {k: (v if k in ("N", "num_grid_points", "M_G") else conv(v)) for k, v in kw.items()}
    cfg = orc.OracleConfig(dtype=T, **kwT)
    Yc = np.ascontiguousarray(Y.T.astype(T))
    Fc = np.ascontiguousarray(F.T.astype(T))
    orc.set_threads(fft_threads, particle_threads)
    try:
        t0 = time.perf_counter()
        V = orc.mobility(cfg, Yc, Fc).T
        dt = time.perf_counter() - t0
    finally:
        orc.set_threads(1, 1)
    return V, dt


def pair_evaluations(Y, L, R_c):
    """(slot, neighbour) pairs within R_c that the correction evaluates (each pair from both of
    its particles), by a periodic k-d tree."""
    from scipy.spatial import cKDTree

    P = np.mod(np.asarray(Y, np.float64).T, L)
    tree = cKDTree(P, boxsize=L)
    return int(tree.count_neighbors(tree, R_c)) - P.shape[0]


def measured_traffic(stage, N, M):
    """HBM bytes per launch of a stage from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary of the same workload (scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_N{N}_grid{M}.json")
    try:
        with open(path) as f:
            return json.load(f)["stages"][stage]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def committed_counters(name):
    """Per-kernel counter rates (VALU, MFMA, atomics) from a committed rocprofv3 summary."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def run_ensemble(steps, warmup):
    """The batched ensemble: 64 suspensions (N log-uniform in [1e3, 5e4], 5-smooth grids 64..240,
    phi = 0.10, h ~ 0.55) with 16 right-hand sides each, as one graph-captured batch apply,
    against the same work as sequential plain device-resident applies (one per system and
    right-hand side), which run first."""
    import torch

    from ForceCouplingMethod import FFCMBatch, FFCMConfig, mobility_device_
    from ForceCouplingMethod.suspensions import ensemble_workload

    R = 16
    ens = [s.typed(np.float32) for s in ensemble_workload(n_rhs=R)]
    total = sum(kw["N"] for kw, _, _ in ens) * R

    def dev(a):
        return torch.from_numpy(np.ravel(a, order="F").copy()).cuda()

    Yd = [dev(Y) for _, Y, _ in ens]
    Fd = [dev(F) for _, _, F in ens]
    # sequential plain applies
    cfgs = [FFCMConfig(gpu_acceleration=True, **kw) for kw, _, _ in ens]
    Vs = [torch.empty_like(f) for f in Fd]

    def sequential():
        for k, cfg in enumerate(cfgs):
            n3 = 3 * ens[k][0]["N"]
            for r in range(R):
                mobility_device_(Vs[k][r * n3:(r + 1) * n3], cfg, Yd[k], Fd[k][r * n3:(r + 1) * n3], synchronize=False)

    sequential()
    torch.cuda.synchronize()
    reps = 2
    t0 = time.perf_counter()
    for _ in range(reps):
        sequential()
    torch.cuda.synchronize()
    ts = (time.perf_counter() - t0) / reps
    Vs = [v.cpu().numpy() for v in Vs]
    for cfg in cfgs:
        cfg.close()
    log(f"ensemble: sequential plain applies {ts * 1e3:.1f} ms")
    # the batch: the first apply times the candidate executions and keeps the fastest
    Vd = [torch.empty_like(f) for f in Fd]
    batch = FFCMBatch([kw for kw, _, _ in ens], n_rhs=R)
    for _ in range(max(1, warmup)):
        batch.mobility_device_(Vd, Yd, Fd, synchronize=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.mobility_device_(Vd, Yd, Fd, synchronize=False)
    torch.cuda.synchronize()
    tb = (time.perf_counter() - t0) / steps
    info = batch.info()
    Vb = [v.cpu().numpy() for v in Vd]
    batch.close()
    log(f"ensemble: batch {tb * 1e3:.1f} ms ({info})")
    worst = max(rel_l2(Vb[k], Vs[k]) for k in range(len(ens)))
    return {"workload": "64 suspensions, N log-uniform in [1e3, 5e4] (sum %d), 5-smooth grids 64..240, phi=0.10, "
                        "16 right-hand sides each, M_G=8, Sigma/sigma=2, R_c=4, Float32, device-resident" % (total // R),
            "particle_rhs_per_s": total / tb, "ms_per_ensemble_apply": tb * 1e3,
            "sequential_plain": {"particle_rhs_per_s": total / ts, "ms": ts * 1e3},
            "speedup_vs_sequential": ts / tb, "rel_l2_vs_sequential_max": worst, "batch": info}


def ensemble_child(steps, timeout=900):
    """run_ensemble in a child process (bench.py --only ensemble): a failure there is reported in
    the line instead of ending the benchmark."""
    import subprocess

    cmd = [sys.executable, "-X", "faulthandler", "-u", os.path.abspath(__file__), "--only", "ensemble",
           "--steps", str(steps)]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    except subprocess.TimeoutExpired:
        return {"error": f"timed out after {timeout} s"}
    sys.stderr.write(r.stderr)
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": f"child exited with {r.returncode}", "stderr_tail": r.stderr[-2000:]}
    return json.loads(lines[-1])


def cufcm_leaf(steps, warmup, threads):
    """The one configuration with published GPU figures (benchmark/cufcm-baseline.md:17-69):
    L = 250, 256^3 (h = 0.9766), M_G = 10, R_c = 5.6419, Sigma/sigma = 2, N = 186,510 (phi = 5 %),
    uniform positions and forces uniform in [-0.5, 0.5]. Context only: the published 20.196 ms
    (ForceCouplingMethod.jl CUDA on an RTX 2080 Ti, host-staged) is other hardware."""
    from ForceCouplingMethod.suspensions import uniform

    T = np.float32
    N, L, M = 186510, 250.0, 256
    Y = uniform(3 * N, 20260713, 0.0, L).reshape(N, 3).T
    F = uniform(3 * N, 20260714, -0.5, 0.5).reshape(N, 3).T
    kw = dict(L=(T(L),) * 3, R_c=T(5.6419), N=N, kernel_widths_ratio=T(2), num_grid_points=(M,) * 3, M_G=10,
              viscosity=T(1))
    Y32, F32 = np.asfortranarray(Y.astype(T)), np.asfortranarray(F.astype(T))
    r = run_plain(kw, Y32, F32, steps, warmup, "cuFCM configuration")
    Vo, _ = oracle_apply(kw, Y32, F32, np.float32, threads, threads)
    out = {"workload": "L=250, 256^3, M_G=10, R_c=5.6419, Sigma/sigma=2, N=186,510 uniform, forces U(-0.5, 0.5), "
                       "Float32 (benchmark/cufcm-baseline.md:17-69)",
           "particles_per_s": N / r["s_per_apply"], "ms_per_apply": r["s_per_apply"] * 1e3,
           "ms_per_apply_host_staged": r["s_per_apply_staged"] * 1e3, "stages": stage_table(r, M),
           "rel_l2_vs_cpu_f32": rel_l2(r["V"], Vo),
           "published_context": {"ms_host_staged": 20.196, "hardware": "RTX 2080 Ti (sm_75), ForceCouplingMethod.jl "
                                 "CUDA mobility!, min of BenchmarkTools samples", "source": "benchmark/cufcm-baseline.md:99-104"}}
    r["cfg"].close()
    return out


def pullback_times(r):
    """hipEvent-timed pullbacks (dF, dY) of the apply in r for an upstream gradient G ~ N(0, 1)."""
    import torch

    from ForceCouplingMethod import mobility_vjp_device_

    dY, dF, dV = r["dev"]
    cfg = r["cfg"]
    dG = torch.from_numpy(np.random.default_rng(0xAD71).normal(size=(r["N"], 3)).astype(np.float32)).cuda()
    gF, gY = torch.empty_like(dF), torch.empty_like(dY)
    vjp = lambda: mobility_vjp_device_(gF, gY, cfg, dY, dF, dG, synchronize=False)
    for _ in range(5):
        vjp()
    evv = event_times(cfg, vjp, 20)
    return {"ms_per_pullback_median": float(np.median(evv)),
            "ratio_to_forward": float(np.median(evv) / np.median(r["event_ms"])),
            "what": "dF = M(Y) G and dY = dL/dY for L = <G, M(Y) F>, cell list and weights recomputed"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU restatement (baseline and error)")
    ap.add_argument("--no-extra", action="store_true", help="only the headline suspension")
    ap.add_argument("--copy-calibration", action="store_true", help="only the known-size copy (PMC scaling)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--only", choices=["headline", "mid", "small", "ensemble", "cufcm"], default=None,
                    help="one leg, for profiling (no CPU work)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs the mobility apply of one suspension on one device "
              "(multi-GPU sharding is a design note, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    from ForceCouplingMethod.suspensions import plain_workload

    threads = host_threads()
    if args.only == "ensemble":
        print(json.dumps(run_ensemble(max(3, args.steps // 4), 1)), flush=True)
        return
    if args.only == "cufcm":
        print(json.dumps(cufcm_leaf(args.steps, args.warmup, threads)), flush=True)
        return
    index = {"mid": 1, "small": 0}.get(args.only, HEADLINE)
    w = plain_workload(index)
    kw, Y, F = w.typed(np.float32)
    N, M = Y.shape[1], kw["num_grid_points"][0]
    log(f"workload {w.name}: N={N} grid={M}^3 phi={w.info['phi']:.3f} cluster fallbacks {w.info['cluster_fallbacks']}")
    r = run_plain(kw, Y, F, args.steps, args.warmup, w.name, cached=True)
    stages = stage_table(r, M)
    dom = max((k for k in stages if "GB_per_s" in stages[k]), key=lambda k: stages[k]["ms"])
    ds = stages[dom]
    out = {
        "metric": METRIC,
        "value": N / r["s_per_apply"],
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["s_per_apply"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": ("plain FFCM, N=1,048,576 spheres, L=280, grid 512^3, phi=0.20 (half uniform RSA, half in "
                                "64 Gaussian clusters of std 8 with RSA rejection, 10,000 draws before a uniform fallback), "
                                "M_G=8, Sigma/sigma=2, R_c=4, forces N(0,1) mean-subtracted, device-resident Y/F/V")
                   if index == HEADLINE else "plain FFCM, N=65,536 spheres, L=140, grid 256^3, RSA phi=0.10",
                   "N": N, "grid": [M, M, M], "M_G": 8, "R_c": 4.0, "L": float(kw["L"][0]),
                   "cluster_fallbacks": w.info["cluster_fallbacks"], "seed": w.info["seed"],
                   "inputs_sha256": {"Y": sha256(Y), "F": sha256(F)}},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": ds["GB_per_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ds["hbm_frac"], "traffic": measured_traffic(dom, N, M)},
        "ms_per_apply_event": {"min": float(np.min(r["event_ms"])), "median": float(np.median(r["event_ms"])),
                               "applies": len(r["event_ms"])},
        "stages": stages,
        "ms_per_apply_host_staged": r["s_per_apply_staged"] * 1e3,
        "host_staged_vs_device_rel_l2": r["staged_vs_device_rel_l2"],
        "cached_positions_apply": r["cached"],
        "graph_apply": r["graph"],
        "launch_shapes": r["info"],
    }
    if args.only is None:
        if not args.no_cpu:
            try:
                evals = pair_evaluations(Y, float(kw["L"][0]), float(kw["R_c"]))
                out["stages"]["correct"]["pair_evaluations"] = evals
                out["stages"]["correct"]["pairs_per_s"] = evals / (out["stages"]["correct"]["ms"] * 1e-3)
            except Exception as e:  # scipy missing: the rate is optional
                log(f"pair count skipped: {e}")
        counters = committed_counters(f"counters_N{N}_grid{M}.json")
        if counters:
            out["counters"] = counters
    if args.only in (None, "headline"):
        out["adjoint"] = pullback_times(r)
    r["cfg"].close()
    del r["dev"]
    if args.only is None and not args.no_cpu:
        # the reference's CPU path restated: FFT threaded over the host cores (fft_threads),
        # the particle loops serial, as src/config.jl:173-176 and src/spread_forces.jl:100 run them
        Vc, tc = oracle_apply(kw, Y, F, np.float32, threads, 1)
        log(f"CPU restatement Float32 ({threads} FFT threads): {tc:.2f} s")
        V64, t64 = oracle_apply(kw, Y, F, np.float64, threads, threads)
        log(f"CPU restatement Float64 ({threads} threads): {t64:.2f} s")
        out["rel_l2_vs_cpu_f32"] = rel_l2(r["V"], Vc)
        out["rel_l2_vs_cpu_f64"] = rel_l2(r["V"], V64)
        out["rel_l2_cpu_f32_vs_f64"] = rel_l2(Vc, V64)
        out["cpu_baseline"] = {"value": N / tc, "unit": "particles/s", "cores": threads, "kind": "port",
                               "sample": f"one full Float32 apply of the same suspension by the C++ restatement of the "
                                         f"reference (oracle/), FFT over {threads} threads and the particle loops on "
                                         f"one, as the reference runs with fft_threads={threads}: {tc:.2f} s; "
                                         f"host CPU {cpu_model()}"}
        del Vc, V64
    if args.only is None and not args.no_extra:
        w1 = plain_workload(1)
        kw1, Y1, F1 = w1.typed(np.float32)
        r1 = run_plain(kw1, Y1, F1, args.steps, args.warmup, w1.name, cached=True)
        mid = {"workload": "plain FFCM, N=65,536, L=140, grid 256^3, RSA phi=0.10",
               "particles_per_s": r1["N"] / r1["s_per_apply"], "ms_per_apply": r1["s_per_apply"] * 1e3,
               "ms_per_apply_event_median": float(np.median(r1["event_ms"])),
               "ms_per_apply_host_staged": r1["s_per_apply_staged"] * 1e3, "stages": stage_table(r1, 256),
               "cached_positions_apply": r1["cached"], "graph_apply": r1["graph"]}
        mid["adjoint"] = pullback_times(r1)
        cfg1 = r1["cfg"]
        if not args.no_cpu:
            Vc1, _ = oracle_apply(kw1, Y1, F1, np.float32, threads, threads)
            mid["rel_l2_vs_cpu_f32"] = rel_l2(r1["V"], Vc1)
            _, t1 = oracle_apply(kw1, Y1, F1, np.float32, 1, 1)
            mid["cpu_1core"] = {"value": r1["N"] / t1, "unit": "particles/s", "cores": 1, "kind": "port",
                                "sample": f"one full Float32 apply of this suspension, every loop on one core: {t1:.2f} s"}
        cfg1.close()
        del r1
        out["mid_size"] = mid
        # the 512-particle / 64^3 suspension, where the host's launches, not HBM, bound an apply:
        # the case a captured apply (ffcm_set_graph) is for
        w0 = plain_workload(0)
        kw0, Y0, F0 = w0.typed(np.float32)
        r0 = run_plain(kw0, Y0, F0, max(args.steps, 50), args.warmup, w0.name, cached=True)
        out["small_size"] = {"workload": "plain FFCM, N=512, L=32, grid 64^3, RSA",
                             "ms_per_apply": r0["s_per_apply"] * 1e3,
                             "ms_per_apply_event_median": float(np.median(r0["event_ms"])),
                             "cached_positions_apply": r0["cached"], "graph_apply": r0["graph"]}
        r0["cfg"].close()
        del r0
        if not args.no_ensemble:
            out["batched"] = ensemble_child(args.steps)
        out["cufcm_config"] = cufcm_leaf(args.steps, min(args.warmup, 20), threads)
    print(json.dumps(out), flush=True)


def copy_calibration():
    """A plain device copy of known size (torch's copy kernel, 1.5 GiB read + 1.5 GiB written),
    run under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to scale those counters to bytes."""
    import torch

    n = 3 * 2 ** 27  # float32 elements: 1.5 GiB
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 3}), flush=True)


if __name__ == "__main__":
    main()
