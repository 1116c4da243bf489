"""Benchmark of the FFCM mobility apply on one MI355X.

One step is one device-resident mobility apply (wrap, hash, radix sort, gather, spread,
FFT Stokes solve, interpolation, correction) over the plain suspension N = 65,536 spheres,
L = 140, grid 256^3, M_G = 8, Sigma/sigma = 2, R_c = 4 (Float32), with positions and forces
already in HBM. Prints one JSON line: particles/s, ms per apply, per-stage device time from
HIP events recorded on the library stream, the HBM roofline of the dominant stage, the
accuracy against the CPU Float32 restatement, and that restatement's own throughput on the
host (one core). The 2^20-particle / 512^3 suspension is measured too and reported under
"large".

    python bench.py [--steps K] [--warmup W] [--no-large] [--no-cpu]
"""
from __future__ import annotations

import argparse
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


def algorithmic_bytes(M, N, C, key_passes, elem=4):
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


def run_gpu(index, steps, warmup, adjoint=False):
    import torch

    from ForceCouplingMethod import FFCMConfig, mobility_, mobility_device_, mobility_vjp_device_
    from ForceCouplingMethod.suspensions import plain_workload

    w = plain_workload(index)
    kw, Y, F = w.typed(np.float32)
    N = Y.shape[1]
    print(f"[bench] workload {w.name}: N={N} grid={kw['num_grid_points']} phi={w.info['phi']:.3f}", flush=True)
    cfg = FFCMConfig(gpu_acceleration=True, **kw)
    V_host = mobility_(np.empty((3, N), np.float32, order="F"), cfg, Y, F)
    dY = torch.from_numpy(np.ascontiguousarray(Y.T)).cuda()
    dF = torch.from_numpy(np.ascontiguousarray(F.T)).cuda()
    dV = torch.empty_like(dY)
    torch.cuda.synchronize()
    for _ in range(warmup):
        mobility_device_(dV, cfg, dY, dF, synchronize=False)
    torch.cuda.synchronize()
    # untimed pass with stage events to attribute time per stage
    cfg.set_timing(True)
    for _ in range(steps):
        mobility_device_(dV, cfg, dY, dF, synchronize=False)
    torch.cuda.synchronize()
    st = cfg.timing()
    cfg.set_timing(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        mobility_device_(dV, cfg, dY, dF, synchronize=False)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    V_dev = dV.cpu().numpy().T
    vjp = None
    if adjoint:
        # pullback (dF, dY) of the same apply for an upstream gradient G ~ N(0, 1)
        dG = torch.from_numpy(np.random.default_rng(0xAD71).normal(size=(N, 3)).astype(np.float32)).cuda()
        gF = torch.empty_like(dF)
        gY = torch.empty_like(dY)
        for _ in range(max(2, warmup // 2)):
            mobility_vjp_device_(gF, gY, cfg, dY, dF, dG, synchronize=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            mobility_vjp_device_(gF, gY, cfg, dY, dF, dG, synchronize=False)
        torch.cuda.synchronize()
        tv = (time.perf_counter() - t0) / steps
        vjp = {"ms_per_pullback": round(tv * 1e3, 4), "ratio_to_forward": round(tv / dt, 3),
               "what": "dF = M(Y) G and dY = dL/dY for L = <G, M(Y) F>, cell list and weights recomputed"}
    # host-staged apply (PCIe-inclusive), for reference
    t0 = time.perf_counter()
    reps = max(3, min(steps, 10))
    for _ in range(reps):
        mobility_(V_host, cfg, Y, F)
    staged = (time.perf_counter() - t0) / reps
    stage_ms = {k: v / max(1, st["calls"]) for k, v in st["ms"].items()}
    C = int(np.prod(cfg.num_cells))
    key_bits = max(1, math.ceil(math.log2(C)))
    ab = algorithmic_bytes(kw["num_grid_points"][0], N, C, math.ceil(key_bits / 8))
    stages = {}
    for k, ms in stage_ms.items():
        entry = {"ms": round(ms, 4)}
        if k in ab:
            gbs = ab[k] / (ms * 1e-3) / 1e9
            entry.update({"algorithmic_bytes": ab[k], "GB_per_s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)})
        stages[k] = entry
    info = cfg.info()
    cfg.close()
    return {"workload": w, "kw": kw, "Y": Y, "F": F, "N": N, "s_per_apply": dt, "s_per_apply_staged": staged,
            "stages": stages, "V": V_dev, "info": info, "adjoint": vjp}


def run_ensemble(steps, warmup):
    """The batched ensemble: 64 suspensions (N log-uniform in [1e3, 5e4], 5-smooth grids 64..240,
    phi = 0.10, h ~ 0.55) with 16 right-hand sides each, as one graph-captured batch apply,
    against the same work as sequential plain device-resident applies (one per system and
    right-hand side)."""
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
    Vd = [torch.empty_like(f) for f in Fd]
    batch = FFCMBatch([kw for kw, _, _ in ens], n_rhs=R)
    for _ in range(max(1, warmup)):
        batch.mobility_device_(Vd, Yd, Fd, synchronize=False)  # the first call autotunes and captures
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.mobility_device_(Vd, Yd, Fd, synchronize=False)
    torch.cuda.synchronize()
    tb = (time.perf_counter() - t0) / steps
    info = batch.info()
    Vb = [v.cpu() for v in Vd]
    batch.close()
    # sequential plain applies of the same work
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
    worst = max(rel_l2(Vb[k].numpy(), Vs[k].cpu().numpy()) for k in range(len(ens)))
    for cfg in cfgs:
        cfg.close()
    return {"workload": "64 suspensions, N log-uniform in [1e3, 5e4] (sum %d), 5-smooth grids 64..240, phi=0.10, "
                        "16 right-hand sides each, M_G=8, Sigma/sigma=2, R_c=4, Float32, device-resident" % (total // R),
            "particle_rhs_per_s": total / tb, "ms_per_ensemble_apply": tb * 1e3,
            "sequential_plain": {"particle_rhs_per_s": total / ts, "ms": ts * 1e3},
            "speedup_vs_sequential": ts / tb, "rel_l2_vs_sequential_max": worst, "batch": info}


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


def ensemble_multi_rhs_profile():
    """The ensemble's systems one at a time through the multi-right-hand-side path (no graph,
    synchronised per system), for a kernel trace per stage."""
    import torch

    from ForceCouplingMethod import FFCMConfig, mobility_multi_device_
    from ForceCouplingMethod.suspensions import ensemble_workload

    R = 16
    ens = [s.typed(np.float32) for s in ensemble_workload(n_rhs=R)]
    runs = []
    for kw, Y, F in ens:
        cfg = FFCMConfig(gpu_acceleration=True, max_rhs=R, **kw)
        Yd = torch.from_numpy(np.ravel(Y, order="F").copy()).cuda()
        Fd = torch.from_numpy(np.ravel(F, order="F").copy()).cuda()
        runs.append((cfg, Yd, Fd, torch.empty_like(Fd)))
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for cfg, Yd, Fd, Vd in runs:
            mobility_multi_device_(Vd, cfg, Yd, Fd, R, synchronize=True)
        print(json.dumps({"pass": rep, "ms": (time.perf_counter() - t0) * 1e3}), flush=True)


def cpu_baseline(r):
    """The CPU restatement (oracle/, single thread) timed on the same suspension."""
    from oracle import oracle as orc

    kw, Y, F = r["kw"], r["Y"], r["F"]
    cfg = orc.OracleConfig(dtype=np.float32, **kw)
    t0 = time.perf_counter()
    V = orc.mobility(cfg, np.ascontiguousarray(Y.T), np.ascontiguousarray(F.T)).T
    dt = time.perf_counter() - t0
    return V, dt


def cpu_f64(r):
    """The CPU restatement in Float64 on the same inputs (the Float32 positions and forces,
    widened): the error of the GPU Float32 apply against the Float64 path."""
    from oracle import oracle as orc

    def wide(v):
        if isinstance(v, tuple):
            return tuple(float(x) for x in v)
        return float(v) if isinstance(v, (float, np.floating)) else v

    kw = {k: (v if k in ("N", "num_grid_points", "M_G") else wide(v)) for k, v in r["kw"].items()}
    cfg = orc.OracleConfig(dtype=np.float64, **kw)
    Y = np.ascontiguousarray(r["Y"].T.astype(np.float64))
    F = np.ascontiguousarray(r["F"].T.astype(np.float64))
    return orc.mobility(cfg, Y, F).T


def measured_traffic(stage, N, M):
    """HBM bytes per launch of a stage from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary of the same workload (scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_N{N}_grid{M}.json")
    try:
        with open(path) as f:
            return json.load(f)["stages"][stage]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", type=int, default=1)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true", help="only the known-size copy (PMC scaling)")
    ap.add_argument("--ensemble-multi-rhs", action="store_true", help="only the ensemble, system by system")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.ensemble_multi_rhs:
        ensemble_multi_rhs_profile()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the mobility apply of one suspension runs on one device "
              "(multi-GPU sharding is a design note, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    r = run_gpu(args.workload, args.steps, args.warmup, adjoint=True)
    N = r["N"]
    dom = max((k for k in r["stages"] if "GB_per_s" in r["stages"][k]), key=lambda k: r["stages"][k]["ms"])
    ds = r["stages"][dom]
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
        "config": {"workload": "plain FFCM, N=65,536 spheres, L=140, grid 256^3, M_G=8, Sigma/sigma=2, R_c=4, "
                               "RSA phi=0.10, device-resident Y/F/V",
                   "N": N, "grid": list(r["kw"]["num_grid_points"]), "M_G": 8, "R_c": 4.0, "L": 140.0},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": ds["GB_per_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ds["hbm_frac"], "traffic": measured_traffic(dom, N, r["kw"]["num_grid_points"][0])},
        "stages": r["stages"],
        "ms_per_apply_host_staged": r["s_per_apply_staged"] * 1e3,
        "launch_shapes": r["info"],
        "adjoint": r["adjoint"],
    }
    if not args.no_cpu:
        Vc, tc = cpu_baseline(r)
        out["rel_l2_vs_cpu_f32"] = rel_l2(r["V"], Vc)
        out["rel_l2_vs_cpu_f64"] = rel_l2(r["V"], cpu_f64(r))
        out["cpu_baseline"] = {"value": N / tc, "unit": "particles/s", "cores": 1, "kind": "port",
                               "sample": f"one full Float32 apply of the same suspension by the C++ restatement "
                                         f"(oracle/), single thread: {tc:.2f} s"}
    if not args.no_large and args.workload != 2:
        L2 = run_gpu(2, max(5, args.steps // 2), max(3, args.warmup // 2))
        out["large"] = {"workload": "plain FFCM, N=1,048,576, L=280, grid 512^3, phi=0.20, half clustered",
                        "particles_per_s": L2["N"] / L2["s_per_apply"], "ms_per_apply": L2["s_per_apply"] * 1e3,
                        "ms_per_apply_host_staged": L2["s_per_apply_staged"] * 1e3, "stages": L2["stages"],
                        "cluster_fallbacks": L2["workload"].info["cluster_fallbacks"]}
    if not args.no_ensemble and args.workload != 2:
        out["batched"] = run_ensemble(max(3, args.steps // 4), 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
