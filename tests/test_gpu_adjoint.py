"""Parity of the mobility pullback (the rrule of mobility!) on the GPU.

For L = <G, M(Y) F> the pullback returns dF = M(Y) G and dY = dL/dY. The oracle's pullback is
pinned in test_oracle.py by float64 central differences of the reference-restated forward
path; here the GPU kernels are compared with it, and at the full 65,536-particle / 256^3 size
with float64 central differences of the GPU's own forward apply on 64 seeded particles.
"""
import warnings

import numpy as np
import pytest

from helpers import REL_L2_F32, rel_l2
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _f64_kw(kw):
    return {k: (tuple(np.float64(x) for x in v) if k == "L" else (np.float64(v) if isinstance(v, np.floating) else v))
            for k, v in kw.items()}


def _config(kw):
    from ForceCouplingMethod import FFCMConfig

    with warnings.catch_warnings():
        warnings.simplefilter("ignore", UserWarning)  # the Float64 backend warns that it is slow
        return FFCMConfig(gpu_acceleration=True, **kw)


def _oracle_vjp(T, kw, Y, F, G):
    cfg = orc.OracleConfig(dtype=T, **kw)
    dF, dY = orc.mobility_vjp(cfg, *(np.ascontiguousarray(np.asarray(x, T).T) for x in (Y, F, G)))
    return dF.T, dY.T


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_vjp_workload_0_matches_oracle(T):
    # N = 512, L = 32, 64^3, M_G = 8, ratio 2, R_c = 4; G i.i.d. N(0, 1)
    from ForceCouplingMethod import mobility_, mobility_vjp_
    from ForceCouplingMethod.suspensions import plain_workload

    w = plain_workload(0)
    kw, Y, F = w.typed(T)
    N = Y.shape[1]
    G = np.asfortranarray(np.random.default_rng(0xAD70).normal(size=(3, N)).astype(T))
    cfg = _config(kw)
    dF = np.empty((3, N), T, order="F")
    dY = np.empty((3, N), T, order="F")
    mobility_vjp_(dF, dY, cfg, Y, F, G)
    oF, oY = _oracle_vjp(T, kw, Y, F, G)
    o64F, o64Y = _oracle_vjp(np.float64, _f64_kw(kw), Y.astype(np.float64), F.astype(np.float64),
                             G.astype(np.float64))
    eF, eY = rel_l2(dF, oF), rel_l2(dY, oY)
    print(f"{np.dtype(T).name}: dF rel L2 {eF:.3e} (vs f64 {rel_l2(dF, o64F):.3e}), "
          f"dY rel L2 {eY:.3e} (vs f64 {rel_l2(dY, o64Y):.3e})")
    if T is np.float32:
        assert eF <= REL_L2_F32
        assert eY <= REL_L2_F32
        assert rel_l2(dY, o64Y) <= REL_L2_F32
    else:
        assert eF <= 1e-11
        assert eY <= 1e-10
    # dF is the forward apply with G as the force
    V = mobility_(np.empty((3, N), T, order="F"), cfg, Y, G)
    assert rel_l2(dF, V) <= (1e-6 if T is np.float32 else 1e-13)


def test_vjp_device_matches_host_and_torch_autograd():
    import torch

    from ForceCouplingMethod import mobility_vjp_, mobility_vjp_device_, rrule_mobility, torch_mobility
    from ForceCouplingMethod.suspensions import plain_workload

    w = plain_workload(0)
    kw, Y, F = w.typed(np.float32)
    N = Y.shape[1]
    G = np.asfortranarray(np.random.default_rng(7).normal(size=(3, N)).astype(np.float32))
    cfg = _config(kw)
    hF = np.empty((3, N), np.float32, order="F")
    hY = np.empty((3, N), np.float32, order="F")
    mobility_vjp_(hF, hY, cfg, Y, F, G)

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a.T).ravel()).to("cuda:0")

    dFd = torch.empty(3 * N, dtype=torch.float32, device="cuda:0")
    dYd = torch.empty_like(dFd)
    torch.cuda.synchronize()
    mobility_vjp_device_(dFd, dYd, cfg, dev(Y), dev(F), dev(G))
    assert rel_l2(dFd.cpu().numpy().reshape(N, 3).T, hF) <= 1e-6
    assert rel_l2(dYd.cpu().numpy().reshape(N, 3).T, hY) <= 1e-6
    # torch.autograd through the same kernels
    Yt = dev(Y).requires_grad_(True)
    Ft = dev(F).requires_grad_(True)
    V = torch_mobility(cfg, Yt, Ft)
    V.backward(dev(G))
    assert rel_l2(Ft.grad.cpu().numpy().reshape(N, 3).T, hF) <= 1e-6
    assert rel_l2(Yt.grad.cpu().numpy().reshape(N, 3).T, hY) <= 1e-6
    # the rrule pair: V and its pullback
    V2, pullback = rrule_mobility(cfg, Y, F)
    assert rel_l2(V.detach().cpu().numpy().reshape(N, 3).T, V2) <= 1e-6
    rY, rF = pullback(G)
    assert rel_l2(rF, hF) <= 1e-6 and rel_l2(rY, hY) <= 1e-6


def _neighbours(Y, n, L, Rmax):
    d = Y - Y[:, n:n + 1]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(axis=0)
    idx = np.nonzero(r2 < Rmax * Rmax)[0]
    return idx[idx != n]


def _smooth(Y, n, i, eps, h, L, Rc):
    """Y[n, i] +- eps moves no stencil anchor and no pair of n across R_c."""
    if any(np.rint((Y[i, n] + s) / h) != np.rint(Y[i, n] / h) for s in (-eps, eps)):
        return False
    for m in _neighbours(Y, n, L, Rc + 1.0):
        rs = []
        for s in (-eps, 0.0, eps):
            d = Y[:, n] - Y[:, m]
            d[i] += s
            d -= L * np.round(d / L)
            rs.append(float(np.sqrt(d @ d)))
        if (min(rs) - Rc) * (max(rs) - Rc) <= 0:
            return False
    return True


def test_vjp_workload_1_central_differences_on_64_particles():
    # N = 65,536, L = 140, 256^3: float64 GPU pullback vs float64 central differences (step
    # 1e-4) of the float64 GPU forward apply on a seeded subset of 64 particles; the float32
    # pullback vs the float64 one over all particles.
    from ForceCouplingMethod import mobility_, mobility_vjp_
    from ForceCouplingMethod.suspensions import plain_workload

    w = plain_workload(1)
    kw32, Y32, F32 = w.typed(np.float32)
    kw = _f64_kw(kw32)
    Y = Y32.astype(np.float64, order="F")
    F = F32.astype(np.float64, order="F")
    N = Y.shape[1]
    G = np.asfortranarray(np.random.default_rng(0xAD71).normal(size=(3, N)))
    cfg = _config(kw)
    dF = np.empty((3, N), order="F")
    dY = np.empty((3, N), order="F")
    mobility_vjp_(dF, dY, cfg, Y, F, G)
    L = float(kw["L"][0])
    h = L / kw["num_grid_points"][0]
    Rc = float(kw["R_c"])
    eps = 1e-4
    scale = float(np.sqrt((dY ** 2).mean()))
    rng = np.random.default_rng(0xFD64)
    V = np.empty((3, N), order="F")
    worst = 0.0
    checked = 0
    for n in rng.choice(N, size=64, replace=False):
        for i in range(3):
            if not _smooth(Y, n, i, eps, h, L, Rc):
                continue
            Lpm = []
            for s in (eps, -eps):
                Yp = Y.copy(order="F")
                Yp[i, n] += s
                mobility_(V, cfg, Yp, F)
                Lpm.append(float((G * V).sum()))
            fd = (Lpm[0] - Lpm[1]) / (2 * eps)
            err = abs(dY[i, n] - fd) / max(abs(fd), scale)
            worst = max(worst, err)
            checked += 1
    print(f"workload 1 FD: {checked} coordinates, worst relative error {worst:.3e}")
    assert checked >= 150
    assert worst <= 1e-6
    # Float32 pullback against the Float64 one
    cfg32 = _config(kw32)
    dF32 = np.empty((3, N), np.float32, order="F")
    dY32 = np.empty((3, N), np.float32, order="F")
    mobility_vjp_(dF32, dY32, cfg32, Y32, F32, G.astype(np.float32))
    e_F, e_Y = rel_l2(dF32, dF), rel_l2(dY32, dY)
    print(f"workload 1 f32 vs f64: dF {e_F:.3e}, dY {e_Y:.3e}")
    assert e_F <= REL_L2_F32
    assert e_Y <= REL_L2_F32


RING_CASES = [
    # (name, L, M, N, FFCM_INTERP_SEGS): ring layouts of the interpolation's column tests
    ("patches_one_layer_blocks", (64.0, 32.0, 24.0), (128, 64, 48), 3000, None),
    ("whole_columns_odd_layers", (16.0, 16.0, 20.0), (32, 32, 40), 900, 1),
    ("three_segments", (24.0, 16.0, 28.0), (48, 32, 56), 2000, 3),
]


def _lattice_positions(rng, L, N, spacing=1.6, jitter=0.2):
    """N distinct sites of a jittered lattice (pair distances >= spacing - 2 sqrt(3) jitter, so
    that no pair sits where the Float32 pair scalars cancel), each moved by a random whole period
    in [-L, 2L) so that stencils wrap around every face."""
    n = [int(l // spacing) for l in L]
    sites = rng.choice(n[0] * n[1] * n[2], size=N, replace=False)
    ix = np.stack([sites % n[0], (sites // n[0]) % n[1], sites // (n[0] * n[1])])
    Y = np.empty((3, N))
    for i in range(3):
        Y[i] = (ix[i] + 0.5) * (L[i] / n[i]) + rng.uniform(-jitter, jitter, N) + L[i] * rng.integers(-1, 2, N)
    return Y


@pytest.mark.parametrize("name,L,M,N,segs", RING_CASES, ids=[c[0] for c in RING_CASES])
def test_vjp_ring_gradient_matches_wave_kernel_and_oracle(name, L, M, N, segs, monkeypatch):
    # The Float32 grid part through the interpolation's LDS-DMA ring (one pass for G . u_F, one
    # adding F . u_G) against the per-particle kernel (FFCM_GRAD_WAVE) and the oracle's pullback.
    from helpers import kwargs

    from ForceCouplingMethod import mobility_vjp_
    from ForceCouplingMethod.suspensions import normal_forces

    T = np.float32
    rng = np.random.default_rng(abs(hash(name)) % 2**32)
    Y = np.asfortranarray(_lattice_positions(rng, L, N).astype(T))
    F = np.asfortranarray(normal_forces(N, 91).astype(T))
    G = np.asfortranarray(rng.normal(size=(3, N)).astype(T))
    kw = kwargs(T, L=L, R_c=2.0, N=N, M=M)
    monkeypatch.setenv("FFCM_INTERP", "dma")
    if segs is None:
        monkeypatch.delenv("FFCM_INTERP_SEGS", raising=False)
    else:
        monkeypatch.setenv("FFCM_INTERP_SEGS", str(segs))
    monkeypatch.delenv("FFCM_GRAD_WAVE", raising=False)
    cfg = _config(kw)
    assert cfg.info()["interp"] == "dma"
    dF = np.empty((3, N), T, order="F")
    dY = np.empty((3, N), T, order="F")
    mobility_vjp_(dF, dY, cfg, Y, F, G)
    monkeypatch.setenv("FFCM_GRAD_WAVE", "1")
    wF = np.empty((3, N), T, order="F")
    wY = np.empty((3, N), T, order="F")
    mobility_vjp_(wF, wY, cfg, Y, F, G)
    cfg.close()
    oF, oY = _oracle_vjp(T, kw, Y, F, G)
    eY, eW = rel_l2(dY, oY), rel_l2(dY, wY)
    print(f"{name}: ring dY vs oracle {eY:.3e}, vs per-particle kernel {eW:.3e}")
    assert eW <= 1e-6
    assert rel_l2(dF, wF) <= 1e-6
    assert eY <= REL_L2_F32
    assert rel_l2(dF, oF) <= REL_L2_F32
