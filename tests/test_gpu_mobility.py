"""End-to-end parity of mobility! on the GPU (test/cuda/test_gpu_mobility.jl:29-55 and the
operator tests of test/accuracy/test_mobility*.jl), plus size-independent properties at the
full benchmark sizes where the oracle would take minutes."""
import numpy as np
import pytest

from helpers import REL_L2_F32, pair, rel_l2
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _oracle_V(T, kw, Y, F):
    cfg = orc.OracleConfig(dtype=T, **kw)
    return orc.mobility(cfg, np.ascontiguousarray(Y.T), np.ascontiguousarray(F.T)).T


def test_plain_workload_0_matches_cpu_float32_and_float64():
    # N = 512, L = 32, 64^3, M_G = 8, ratio 2, R_c = 4: rel L2 <= 1e-5 vs the Float32 CPU path
    from ForceCouplingMethod import FFCMConfig, mobility_
    from ForceCouplingMethod.suspensions import plain_workload

    w = plain_workload(0)
    kw, Y, F = w.typed(np.float32)
    cfg = FFCMConfig(gpu_acceleration=True, **kw)
    V = mobility_(np.empty((3, w.Y.shape[1]), np.float32, order="F"), cfg, Y, F)
    V32 = _oracle_V(np.float32, kw, Y, F)
    kw64 = {k: (tuple(np.float64(x) for x in v) if k == "L" else (np.float64(v) if isinstance(v, np.floating) else v))
            for k, v in kw.items()}
    V64 = _oracle_V(np.float64, kw64, Y.astype(np.float64), F.astype(np.float64))
    e32 = rel_l2(V, V32)
    e64 = rel_l2(V, V64)
    print(f"workload 0: rel L2 vs CPU f32 {e32:.3e}, vs CPU f64 {e64:.3e}, CPU f32 vs f64 {rel_l2(V32, V64):.3e}")
    assert e32 <= REL_L2_F32
    assert e64 <= 2e-5


def test_float64_backend_matches_cpu_float64():
    from ForceCouplingMethod import FFCMConfig, mobility_
    from ForceCouplingMethod.suspensions import plain_workload

    w = plain_workload(0)
    kw, Y, F = w.typed(np.float64)
    with pytest.warns(UserWarning):
        cfg = FFCMConfig(gpu_acceleration=True, **kw)
    V = mobility_(np.empty((3, w.Y.shape[1]), np.float64, order="F"), cfg, Y, F)
    assert rel_l2(V, _oracle_V(np.float64, kw, Y, F)) <= 1e-11


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_random_small_systems_match_oracle(T):
    from ForceCouplingMethod import FFCMConfig, mobility_

    rng = np.random.default_rng(2024)
    for M, Lf, N, MG, ratio, Rc in ((16, 8.0, 40, 8, 2.0, 1.0), (24, 12.0, 300, 6, 1.5, 2.5), (30, 15.0, 700, 8, 2.0, 3.0),
                                    (36, 18.0, 1000, 10, 2.5, 4.0)):
        kw = dict(L=(T(Lf),) * 3, R_c=T(Rc), N=N, kernel_widths_ratio=T(ratio), num_grid_points=(M,) * 3, M_G=MG,
                  viscosity=T(0.7))
        Y = np.asfortranarray(((3 * rng.random((3, N)) - 1) * Lf).astype(T))
        F = np.asfortranarray(rng.normal(size=(3, N)).astype(T))
        cfg = FFCMConfig(gpu_acceleration=True, **kw)
        V = mobility_(np.zeros((3, N), T, order="F"), cfg, Y, F)
        assert rel_l2(V, _oracle_V(T, kw, Y, F)) <= (REL_L2_F32 if T is np.float32 else 1e-11), (M, N)


def _periodic_self_mobility_xx(sigma, mu, L, n_max=15):
    # test/test_utilities.jl:112-123
    n = np.arange(-n_max, n_max + 1)
    kx, ky, kz = np.meshgrid(2 * np.pi * n / L, 2 * np.pi * n / L, 2 * np.pi * n / L, indexing="ij")
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    m = k2 > 0
    return float((np.exp(-sigma ** 2 * k2[m]) / (mu * k2[m]) * (1 - kx[m] ** 2 / k2[m])).sum() / L ** 3)


def test_single_sphere_self_mobility():
    # test/accuracy/test_single_sphere_mobility.jl:6-64: Σ = σ, M = 32, M_G = 21, rtol 1e-5
    from ForceCouplingMethod import FFCMConfig, mobility_

    T = np.float64
    cfg = FFCMConfig(gpu_acceleration=True, T=T, L=(8.0, 8.0, 8.0), R_c=1.0, N=1, kernel_widths_ratio=1.0,
                     num_grid_points=(32, 32, 32), M_G=21, viscosity=1.0)
    V = mobility_(np.zeros((3, 1), T, order="F"), cfg, np.array([[4.0], [4.0], [4.0]]), np.array([[1.0], [0.0], [0.0]]))
    ref = _periodic_self_mobility_xx(float(cfg.sigma), 1.0, 8.0)
    assert V[0, 0] == pytest.approx(ref, rel=1e-5)
    assert abs(V[1, 0]) < 1e-12 and abs(V[2, 0]) < 1e-12


def test_operator_identities_and_errors():
    # test/accuracy/test_mobility.jl:6-198 (linearity, symmetry, SPD, mul!, *, traits, errors)
    from ForceCouplingMethod import ArgumentError, DimensionMismatch, FFCMConfig, FFCMMobility, mobility_, mul_

    T = np.float32
    rng = np.random.default_rng(7)
    N = 200
    cfg = FFCMConfig(gpu_acceleration=True, L=(T(16),) * 3, R_c=T(3), N=N, kernel_widths_ratio=T(2),
                     num_grid_points=(32,) * 3, M_G=8, viscosity=T(1))
    Y = np.asfortranarray(rng.uniform(0, 16, (3, N)).astype(T))
    F = np.asfortranarray(rng.normal(size=(3, N)).astype(T))
    G = np.asfortranarray(rng.normal(size=(3, N)).astype(T))
    Y0 = Y.copy()
    VF = mobility_(np.empty((3, N), T, order="F"), cfg, Y, F)
    VG = mobility_(np.empty((3, N), T, order="F"), cfg, Y, G)
    assert np.array_equal(Y, Y0)  # the caller's Y is never mutated
    Vc = mobility_(np.empty((3, N), T, order="F"), cfg, Y, 2 * F - 3 * G)
    assert rel_l2(Vc, 2 * VF - 3 * VG) <= 1e-5
    assert float((G * VF).sum()) == pytest.approx(float((F * VG).sum()), rel=1e-5)
    assert float((F * VF).sum()) > 0
    M = FFCMMobility(cfg, Y)
    assert M.size() == (3 * N, 3 * N) and M.issymmetric() and M.isposdef()
    v = np.empty(3 * N, T)
    mul_(v, M, F.reshape(-1, order="F"))
    # repeated applies agree to the reordering of the spread's LDS accumulation
    assert rel_l2(v, VF.reshape(-1, order="F")) <= 1e-6
    w = np.full(3 * N, np.nan, T)
    mul_(w, M, F.reshape(-1, order="F"), 2.0, 0.0)  # beta = 0 overwrites NaN
    assert rel_l2(w, 2 * v) <= 1e-6
    assert rel_l2(M @ F, VF) <= 1e-6
    with pytest.raises(DimensionMismatch):
        mobility_(np.empty((3, N - 1), T), cfg, Y, F)
    with pytest.raises(DimensionMismatch):
        mobility_(np.empty((3, N), T), cfg, Y[:, :-1], F)
    with pytest.raises(ArgumentError):
        FFCMMobility(cfg, Y[:, :-1])


def test_device_resident_apply_matches_host_apply():
    import torch

    from ForceCouplingMethod import FFCMConfig, mobility_, mobility_device_
    from ForceCouplingMethod.suspensions import plain_workload

    w = plain_workload(0)
    kw, Y, F = w.typed(np.float32)
    cfg = FFCMConfig(gpu_acceleration=True, **kw)
    V = mobility_(np.empty(Y.shape, np.float32, order="F"), cfg, Y, F)
    dY = torch.from_numpy(np.ascontiguousarray(Y.T)).cuda()
    dF = torch.from_numpy(np.ascontiguousarray(F.T)).cuda()
    dV = torch.empty_like(dY)
    torch.cuda.synchronize()
    mobility_device_(dV, cfg, dY, dF)
    assert rel_l2(dV.cpu().numpy().T, V) <= 1e-6


def test_workload_1_matches_cpu_float32():
    # N = 65,536, L = 140, 256^3 (about 15 s of oracle time)
    from ForceCouplingMethod import FFCMConfig, mobility_
    from ForceCouplingMethod.suspensions import plain_workload

    w = plain_workload(1)
    kw, Y, F = w.typed(np.float32)
    cfg = FFCMConfig(gpu_acceleration=True, **kw)
    V = mobility_(np.empty(Y.shape, np.float32, order="F"), cfg, Y, F)
    e = rel_l2(V, _oracle_V(np.float32, kw, Y, F))
    print(f"workload 1: rel L2 vs CPU f32 {e:.3e}")
    assert e <= REL_L2_F32


def test_workload_2_properties_at_full_size():
    # N = 2^20, 512^3: linearity, symmetry <G, M F> = <F, M G>, positivity, determinism of
    # the atomic-free pipeline, and translation by whole cells (no oracle at this size)
    from ForceCouplingMethod import FFCMConfig, mobility_
    from ForceCouplingMethod.suspensions import normal_forces, plain_workload

    w = plain_workload(2)
    kw, Y, F = w.typed(np.float32)
    N = Y.shape[1]
    G = np.asfortranarray(normal_forces(N, 99).astype(np.float32))
    cfg = FFCMConfig(gpu_acceleration=True, **kw)
    out = lambda: np.empty((3, N), np.float32, order="F")
    VF = mobility_(out(), cfg, Y, F)
    VG = mobility_(out(), cfg, Y, G)
    VF2 = mobility_(out(), cfg, Y, F)
    assert np.isfinite(VF).all()
    # every kernel is atomic-free except the LDS accumulation of the spread; allow its reordering
    assert rel_l2(VF2, VF) <= 1e-6
    d = lambda x, y: float((x.astype(np.float64) * y.astype(np.float64)).sum())
    scale = np.linalg.norm(G.astype(np.float64)) * np.linalg.norm(VF.astype(np.float64))
    assert abs(d(G, VF) - d(F, VG)) <= 1e-5 * scale
    assert d(F, VF) > 0
    Vlin = mobility_(out(), cfg, Y, F + 0.5 * G)
    assert rel_l2(Vlin, VF + 0.5 * VG) <= 1e-5
    # translating every particle by L/2 = 256 h = 35 cells leaves V unchanged up to the
    # rounding of the shifted positions (which can move a rare stencil anchor across a tie)
    Vt = mobility_(out(), cfg, Y + np.float32(140.0), F)
    assert rel_l2(Vt, VF) <= 5e-4
