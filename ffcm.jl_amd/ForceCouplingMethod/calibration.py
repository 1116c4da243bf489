"""Calibration of the method parameters to the hardware (the helpers the reference plans but
does not provide: README.md:80-87, "Parameter optimisation").

The Fast FCM apply has three method parameters, ``kernel_widths_ratio`` (Sigma/sigma), ``R_c``
and ``M_G``, plus the grid spacing h = L / M. They trade grid work (spread, FFT,
interpolation: more grid points and wider stencils for a wider kernel resolved more finely)
against pair work (the correction's cutoff grows with Sigma) at a given accuracy
(docs/src/method/overview.md:29-31). ``calibrate`` picks, for a target relative error, the
parameter set with the lowest predicted apply time on this GPU:

* **Accuracy**, measured with this backend on a sample suspension at the target volume
  fraction. The error of a parameter set is its relative L2 distance to a resolved apply of
  the Sigma = sigma limit (no pair correction, Sigma / h = 2, a stencil of +-5 sigma), which
  is the FCM operator that the modified kernel and the correction reproduce (Su & Keaveny 2024,
  section 5). The grid error (Sigma / h, M_G) and the cutoff error (R_c / Sigma) are measured
  separately, each with the other made negligible, and added; the chosen set is then measured
  as a whole.
* **Cost**, from a per-stage time model of this device fitted to HIP-event stage times of
  probe applies: cell list a N + b; spread and interpolation a M^3 + b N M_G^3 + c; FFT
  a B_fft(M) + b with B_fft = 3 (8 M^3 + 64 H), H = (M/2 + 1) M^2; correction a P + b N + c
  with P = N n (4/3) pi R_c^3 the expected pairs at number density n = N / L^3.

Grid sizes are 5-smooth (the FFT's compile-time plans). Everything runs on the GPU through the
public API; nothing here is a CPU fallback.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .config import FFCMConfig
from .errors import ArgumentError
from .mobility import mobility_device_
from .suspensions import normal_forces, rsa_positions, smooth5_at_least

# Candidate grid. Measured on this backend (phi = 0.1 sample against the resolved operator):
# the error is set by the stencil's half-width in units of Sigma, (M_G / 2 - 1/2) / (Sigma / h),
# not by Sigma / h itself (Sigma / h = 1 with M_G = 12, 5.5 Sigma: 6e-6; Sigma / h = 2 with
# M_G = 16, 3.75 Sigma: 7e-4), and by R_c / Sigma (ratio 1.25: 2e-3 at 3, 4e-5 at 5).
RATIOS = (1.25, 1.5, 2.0, 2.5, 3.0)      # Sigma / sigma
SIGMA_PER_H = (0.75, 1.0, 1.25, 1.5)     # Sigma / h
STENCILS = (6, 8, 10, 12, 14, 16)        # M_G
CUTOFFS = (3.0, 4.0, 5.0, 6.0)           # R_c / Sigma (the correction decays like exp(-R_c^2 / 4 Sigma^2))
SIGMA_UNIT = 1.0 / math.sqrt(math.pi)  # sigma for a = 1 (src/config.jl:338)


def fft_bytes(M: int) -> int:
    H = (M // 2 + 1) * M * M
    return 3 * (8 * M ** 3 + 64 * H)


@dataclass
class Candidate:
    ratio: float
    sigma_per_h: float
    M_G: int
    cutoff: float  # R_c / Sigma

    @property
    def Sigma(self) -> float:
        return self.ratio * SIGMA_UNIT

    def grid(self, L: float) -> int:
        return smooth5_at_least(math.ceil(L * self.sigma_per_h / self.Sigma))

    def R_c(self) -> float:
        return self.cutoff * self.Sigma

    def feasible(self, L: float) -> bool:
        M = self.grid(L)
        return self.M_G <= M and self.R_c() <= L / 2 and M <= 512


@dataclass
class CostModel:
    """Per-stage apply time (ms) of one device as a function of the parameters."""

    coef: dict = field(default_factory=dict)

    def stages(self, N: int, L: float, M: int, M_G: int, R_c: float) -> dict:
        c = self.coef
        n = N / L ** 3
        pairs = N * n * 4.0 / 3.0 * math.pi * R_c ** 3
        return {
            "cell_list": c["cell"][0] * N + c["cell"][1],
            "spread": c["spread"][0] * M ** 3 + c["spread"][1] * N * M_G ** 3 + c["spread"][2],
            "stokes_fft": c["fft"][0] * fft_bytes(M) + c["fft"][1],
            "interpolate": c["interp"][0] * M ** 3 + c["interp"][1] * N * M_G ** 3 + c["interp"][2],
            "correct": c["correct"][0] * pairs + c["correct"][1] * N + c["correct"][2],
        }

    def predict(self, N, L, M, M_G, R_c) -> float:
        return float(sum(max(0.0, v) for v in self.stages(N, L, M, M_G, R_c).values()))


def _suspension(N, L, seed):
    Y, _ = rsa_positions(N, L, seed)
    F = normal_forces(N, seed ^ 0x5EED)
    return Y, F


def _apply(Y, F, L, ratio, M, M_G, R_c, T=np.float32, timing=False, reps=1):
    """One device-resident apply with these parameters; returns (V as float64 3xN, stage ms)."""
    import torch

    N = Y.shape[1]
    cfg = FFCMConfig(L=(T(L),) * 3, R_c=T(R_c), N=N, kernel_widths_ratio=T(ratio), num_grid_points=(M,) * 3,
                     M_G=int(M_G), viscosity=T(1), gpu_acceleration=True)
    try:
        dY = torch.from_numpy(np.ascontiguousarray(Y.T.astype(T))).cuda()
        dF = torch.from_numpy(np.ascontiguousarray(F.T.astype(T))).cuda()
        dV = torch.empty_like(dY)
        mobility_device_(dV, cfg, dY, dF)
        stages = None
        if timing:
            cfg.set_timing(True)
            for _ in range(reps):
                mobility_device_(dV, cfg, dY, dF, synchronize=False)
            cfg.synchronize()
            t = cfg.timing()
            cfg.set_timing(False)
            stages = {k: v / max(1, t["calls"]) for k, v in t["ms"].items()}
        return dV.cpu().numpy().T.astype(np.float64), stages
    finally:
        cfg.close()


def reference_grid(L: float):
    """Grid and stencil of the resolved Sigma = sigma apply in a box of side L: Sigma / h >= 2
    on a multiple of 16 points (whole 16 x 16 x 8 spread tiles, so that the wide stencil's
    tiles fit the spread's limits) and a stencil of +-5 sigma."""
    M = 16 * smooth5_at_least(math.ceil(L * 2.0 / SIGMA_UNIT / 16.0))
    return M, 2 * math.ceil(5.0 * SIGMA_UNIT / (L / M))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def fit_cost_model(T=np.float32, reps=5) -> CostModel:
    """Times probe applies on the current device and fits the per-stage model."""
    probes = []
    # suspensions at phi = 0.12 with grids and stencils varied independently of N
    for N, L, M, MG, Rc in ((8000, 65.0, 64, 6, 2.5), (8000, 65.0, 128, 8, 3.5), (30000, 101.0, 128, 8, 3.0),
                            (30000, 101.0, 180, 10, 4.0), (60000, 128.0, 160, 6, 4.5), (60000, 128.0, 240, 12, 3.0),
                            (100000, 152.0, 256, 8, 4.0)):
        Y, F = _suspension(N, L, 0xCA1 + N + M)
        _, st = _apply(Y, F, L, 2.0, M, MG, Rc, T=T, timing=True, reps=reps)
        probes.append((N, L, M, MG, Rc, st))

    def lstsq(rows, y):
        A = np.asarray(rows, float)
        return [float(v) for v in np.linalg.lstsq(A, np.asarray(y, float), rcond=None)[0]]

    coef = {
        "cell": lstsq([[N, 1] for N, L, M, MG, Rc, st in probes], [st["cell_list"] for *_, st in probes]),
        "spread": lstsq([[M ** 3, N * MG ** 3, 1] for N, L, M, MG, Rc, st in probes], [st["spread"] for *_, st in probes]),
        "fft": lstsq([[fft_bytes(M), 1] for N, L, M, MG, Rc, st in probes], [st["stokes_fft"] for *_, st in probes]),
        "interp": lstsq([[M ** 3, N * MG ** 3, 1] for N, L, M, MG, Rc, st in probes],
                        [st["interpolate"] for *_, st in probes]),
        "correct": lstsq([[N * (N / L ** 3) * 4 / 3 * math.pi * Rc ** 3, N, 1] for N, L, M, MG, Rc, st in probes],
                         [st["correct"] for *_, st in probes]),
    }
    return CostModel(coef)


def select(candidates, error_of, cost_of, tol):
    """The feasible candidate of lowest cost whose error is at most tol (None if none is)."""
    best = None
    for c in candidates:
        e = error_of(c)
        if e is None or not e <= tol:
            continue
        t = cost_of(c)
        if best is None or t < best[1]:
            best = (c, t, e)
    return best


def calibrate(L: float, N: int, tol: float = 1e-4, cost_model: CostModel | None = None, sample_N: int = 300,
              seed: int = 0xCA11, T=np.float32) -> dict:
    """The cheapest (kernel_widths_ratio, M_G, grid, R_c) for a cubic box of side L with N unit
    spheres whose velocities must be within relative L2 ``tol`` of the FCM operator.

    Returns the FFCMConfig keyword arguments of the choice, its predicted per-stage times on
    this device, the errors measured on the sample and every evaluated candidate."""
    phi = N * 4.0 / 3.0 * math.pi / L ** 3
    model = cost_model or fit_cost_model(T)
    # sample box at the target volume fraction, wide enough for the largest cutoff
    Rmax = (max(CUTOFFS) + 1.0) * max(RATIOS) * SIGMA_UNIT
    Ls = max(2.0 * Rmax + 1.0, (sample_N * 4.0 / 3.0 * math.pi / max(phi, 1e-3)) ** (1.0 / 3.0))
    Ns = max(16, min(sample_N, int(round(phi * Ls ** 3 / (4.0 / 3.0 * math.pi)))))
    Y, F = _suspension(Ns, Ls, seed)
    # the resolved Sigma = sigma operator: no correction, Sigma / h = 2 (aliasing below 1e-8), a
    # stencil of +-5 sigma (truncated mass below 1e-6)
    M_ref, MG_ref = reference_grid(Ls)
    V_ref, _ = _apply(Y, F, Ls, 1.0, M_ref, MG_ref, min(2.0, Ls / 2), T=np.float64)
    grid_err, cut_err = {}, {}
    for ratio in RATIOS:
        # the grid errors with a cutoff past the largest candidate, the cutoff errors on a grid
        # and stencil whose error is negligible (Sigma / h = 1, half-width 7.5 Sigma)
        cut_big = Candidate(ratio, 1.0, 16, min(max(CUTOFFS) + 1.0, Ls / 2 / (ratio * SIGMA_UNIT)))
        for sh in SIGMA_PER_H:
            for MG in STENCILS:
                c = Candidate(ratio, sh, MG, cut_big.cutoff)
                if not c.feasible(Ls):
                    continue
                try:
                    V, _ = _apply(Y, F, Ls, ratio, c.grid(Ls), MG, c.R_c(), T=T)
                except ArgumentError:  # a grid this backend does not tile or transform
                    continue
                grid_err[(ratio, sh, MG)] = _rel(V, V_ref)
        for cut in CUTOFFS:
            c = Candidate(ratio, 1.0, 16, cut)
            if not c.feasible(Ls):
                continue
            try:
                V, _ = _apply(Y, F, Ls, ratio, c.grid(Ls), 16, c.R_c(), T=T)
            except ArgumentError:
                continue
            cut_err[(ratio, cut)] = _rel(V, V_ref)
    cands = [Candidate(r, sh, mg, cut) for r in RATIOS for sh in SIGMA_PER_H for mg in STENCILS for cut in CUTOFFS]
    cands = [c for c in cands if c.feasible(L) and (c.ratio, c.sigma_per_h, c.M_G) in grid_err
             and (c.ratio, c.cutoff) in cut_err]

    def error_of(c):
        return grid_err[(c.ratio, c.sigma_per_h, c.M_G)] + cut_err[(c.ratio, c.cutoff)]

    def cost_of(c):
        return model.predict(N, L, c.grid(L), c.M_G, c.R_c())

    best = select(cands, error_of, cost_of, tol)
    if best is None:
        raise ValueError(f"no candidate meets a relative error of {tol:g} (smallest: "
                         f"{min(error_of(c) for c in cands):.2e})")
    c, t, e = best
    V, _ = _apply(Y, F, Ls, c.ratio, c.grid(Ls), c.M_G, c.R_c(), T=T)
    return {
        "config": dict(L=(L, L, L), R_c=c.R_c(), N=N, kernel_widths_ratio=c.ratio, num_grid_points=(c.grid(L),) * 3,
                       M_G=c.M_G, viscosity=1.0),
        "sigma_per_h": c.sigma_per_h, "cutoff_per_Sigma": c.cutoff,
        "predicted_ms": t, "predicted_stages_ms": model.stages(N, L, c.grid(L), c.M_G, c.R_c()),
        "sample_error_estimate": e, "sample_error_measured": _rel(V, V_ref),
        "sample": {"N": Ns, "L": Ls, "reference": {"ratio": 1.0, "grid": M_ref, "M_G": MG_ref}},
        "candidates": [{"ratio": x.ratio, "sigma_per_h": x.sigma_per_h, "M_G": x.M_G, "cutoff_per_Sigma": x.cutoff,
                        "error": error_of(x), "predicted_ms": cost_of(x)} for x in cands],
    }
