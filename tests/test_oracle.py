"""Pins the CPU restatement (oracle/) before it is trusted as the parity oracle.

Every check here is independent of the oracle's own formulas: the reference test suite's
known answers (restated with their file:line), closed forms evaluated in float64 numpy,
numpy.fft, brute force on tiny inputs, and operator identities (adjointness, symmetry).
"""
import itertools

import numpy as np
import pytest
from scipy.special import erf

from oracle import oracle as orc

TYPES = [np.float32, np.float64]


def rtol_of(T):
    return float(np.sqrt(np.finfo(T).eps))


def atol_of(T):
    # test/test_utilities.jl:8-9
    return 1e-6 if T is np.float32 else 1e-10


def grid_kwargs(L):
    """test/test_utilities.jl:17-30 (_fcm_grid_kwargs)."""
    h_target = L[0] / 8
    M = tuple(int(round(L[i] / h_target)) for i in range(3))
    return dict(kernel_widths_ratio=2.0, num_grid_points=M, M_G=min(8, min(M)), viscosity=1.0)


# ----------------------------------------------------------------------------- derived scalars
def test_cell_geometry_known_answers():
    # test/accuracy/test_cell_geometry.jl:4-35
    cfg = orc.OracleConfig(L=(4.0, 6.0, 8.0), R_c=1.0, N=10, dtype=np.float64, **grid_kwargs((4.0, 6.0, 8.0)))
    assert cfg.num_cells == (4, 6, 8)
    for i, L in enumerate((4.0, 6.0, 8.0)):
        assert cfg.cell_size[i] == pytest.approx(L / cfg.num_cells[i])
        assert cfg.inv_cell_size[i] == pytest.approx(1 / cfg.cell_size[i])
    cfg = orc.OracleConfig(L=(2.0, 2.0, 2.0), R_c=1.0, N=1, dtype=np.float64, **grid_kwargs((2.0, 2.0, 2.0)))
    assert cfg.num_cells == (3, 3, 3)


def test_fcm_grid_known_answers():
    # test/accuracy/test_fcm_grid.jl:29-69
    base = dict(L=(4.0, 4.0, 4.0), R_c=1.0, N=10, kernel_widths_ratio=2.0, num_grid_points=(8, 8, 8), M_G=8, viscosity=1.0)
    cfg = orc.OracleConfig(dtype=np.float64, **base)
    assert cfg.sigma == pytest.approx(1 / np.sqrt(np.pi))
    assert cfg.h == pytest.approx(0.5) and cfg.inv_h == pytest.approx(2.0)
    cfg = orc.OracleConfig(dtype=np.float64, **{**base, "kernel_widths_ratio": 2.5})
    assert cfg.Sigma == pytest.approx(2.5 * cfg.sigma)
    with pytest.raises(ValueError):
        orc.OracleConfig(dtype=np.float64, **{**base, "L": (4.0, 6.0, 8.0)})
    cfg = orc.OracleConfig(dtype=np.float64, **{**base, "L": (4.0, 8.0, 12.0), "num_grid_points": (8, 16, 24)})
    assert cfg.h == pytest.approx(0.5)


def test_wavevector_layout():
    # test/accuracy/test_stokes_solve.jl:66-92: k_x non-negative half; k_y, k_z in FFTW wrap order
    # with index M/2+1 holding +M/2.
    for M in (8, 7):
        cfg = orc.OracleConfig(L=(4.0, 4.0, 4.0), R_c=1.0, N=1, kernel_widths_ratio=2.0, num_grid_points=(M, M, M),
                               M_G=min(M, 6), viscosity=1.0, dtype=np.float64)
        m = np.fft.fftfreq(M, d=1.0 / M)
        if M % 2 == 0:
            m[M // 2] = M // 2
        np.testing.assert_allclose(cfg.k_y, 2 * np.pi * m / 4.0, rtol=1e-15, atol=1e-15)
        np.testing.assert_allclose(cfg.k_x, 2 * np.pi * np.arange(M // 2 + 1) / 4.0, rtol=1e-15)


# ----------------------------------------------------------------------------- wrap / hash / sort
def _cfg_small(T, N, L=(4.0, 6.0, 8.0)):
    return orc.OracleConfig(L=L, R_c=1.0, N=N, dtype=T, **grid_kwargs(L))


@pytest.mark.parametrize("T", TYPES)
def test_wrap_known_answers(T):
    # test/accuracy/test_wrap_positions.jl:9-62
    cfg = _cfg_small(T, 1)
    assert np.all(orc.wrap_positions(cfg, [[4.0, 6.0, 8.0]]) == 0)
    Y = np.array([[1.5, 2.5, 3.5], [0.0, 5.999, 0.0], [3.999, 0.0, 7.999]], dtype=T)
    cfg = _cfg_small(T, 3)
    assert np.array_equal(orc.wrap_positions(cfg, Y), Y)
    Y = np.array([[-1.3, 0.0, 8.5], [9.7, 7.5, -0.1], [-100.0, 20.0, -8.0]], dtype=T)
    once = orc.wrap_positions(cfg, Y)
    assert np.array_equal(orc.wrap_positions(cfg, once), once)
    assert np.all(once >= 0) and np.all(once < np.array([4, 6, 8], dtype=T))
    eps = T(1e-9) if T is np.float64 else T(1e-6)
    cfg1 = _cfg_small(T, 1)
    w = orc.wrap_positions(cfg1, np.full((1, 3), -eps, dtype=T))
    np.testing.assert_allclose(w[0], np.array([4, 6, 8], dtype=np.float64) - float(eps), rtol=1e-7)
    # Julia mod semantics: matches numpy's floor-mod for these values.
    rng = np.random.default_rng(7)
    Y = (rng.uniform(-3, 4, size=(200, 3)) * np.array([4, 6, 8])).astype(T)
    cfg = _cfg_small(T, 200)
    np.testing.assert_array_equal(orc.wrap_positions(cfg, Y), np.mod(Y, np.array([4, 6, 8], dtype=T)))


def test_hash_known_answers():
    # test/accuracy/test_assign_cells_kernel.jl:5-63 (inv_cell_size 1, cells 4^3 or (4,6,8))
    cfg = orc.OracleConfig(L=(4.0, 4.0, 4.0), R_c=1.0, N=1, dtype=np.float64, **grid_kwargs((4.0, 4.0, 4.0)))
    assert cfg.inv_cell_size == (1.0, 1.0, 1.0) and cfg.num_cells == (4, 4, 4)
    assert orc.assign_cells(cfg, [[0.5, 0.5, 0.5]])[0] == 0
    assert orc.assign_cells(cfg, [[2.5, 2.5, 2.5]])[0] == 42
    assert orc.assign_cells(cfg, [[4.0, 4.0, 4.0]])[0] == 63  # clamp
    cfg = orc.OracleConfig(L=(4.0, 6.0, 8.0), R_c=1.0, N=1, dtype=np.float64, **grid_kwargs((4.0, 6.0, 8.0)))
    assert orc.assign_cells(cfg, [[0.5, 2.5, 5.5]])[0] == 128


def test_counting_sort_known_answers():
    # test/accuracy/test_sort_particles_by_cell.jl:7-72
    orig, cs, ce = orc.build_cell_list(np.array([2, 0, 2, 1]), 3)
    h = np.array([2, 0, 2, 1])
    assert list(h[orig - 1]) == sorted(h[orig - 1]) and sorted(orig) == [1, 2, 3, 4]
    orig, cs, ce = orc.build_cell_list(np.array([0, 2, 0, 2]), 3)
    assert ce[1] < cs[1]
    orig, cs, ce = orc.build_cell_list(np.array([0, 1, 0, 1, 0]), 2)
    assert list(orig[cs[0] - 1:ce[0]]) == [1, 3, 5]
    assert list(orig[cs[1] - 1:ce[1]]) == [2, 4]


def test_counting_sort_is_stable_argsort():
    rng = np.random.default_rng(3)
    for N, C in ((1, 1), (50, 7), (5000, 343), (3000, 10000)):
        h = rng.integers(0, C, size=N).astype(np.int32)
        orig, cs, ce = orc.build_cell_list(h, C)
        np.testing.assert_array_equal(orig - 1, np.argsort(h, kind="stable"))
        counts = np.bincount(h, minlength=C)
        np.testing.assert_array_equal(cs, np.concatenate([[0], np.cumsum(counts)[:-1]]) + 1)
        np.testing.assert_array_equal(ce, np.cumsum(counts))


@pytest.mark.parametrize("T", TYPES)
def test_gather(T):
    # test/accuracy/test_sort_particles_by_cell.jl:74-87
    rng = np.random.default_rng(1)
    Y = rng.random((6, 3)).astype(T)
    F = rng.random((6, 3)).astype(T)
    orig = np.array([4, 1, 6, 2, 5, 3], dtype=np.int32)
    cfg = _cfg_small(T, 6)
    Ys, Fs = orc.gather(cfg, orig, Y, F)
    assert np.array_equal(Ys, Y[orig - 1]) and np.array_equal(Fs, F[orig - 1])


# ----------------------------------------------------------------------------- stencil + spread
@pytest.mark.parametrize("T", TYPES)
def test_kernel_coefficients(T):
    sigma = T(1) / np.sqrt(T(np.pi))
    a0, a2, inv_norm, inv_2S2 = orc.kernel_coefficients(sigma, sigma, T)
    assert a0 == 1 and a2 == 0  # test_modified_kernel_coefficients.jl: Σ = σ exactly
    a0, a2, _, _ = orc.kernel_coefficients(sigma, T(2) * sigma, T)
    assert a0 == pytest.approx(2.125, rel=1e-6) and a2 == pytest.approx(-0.294524, rel=1e-5)


def modified_kernel_f64(r2, sigma, Sigma):
    """Δ̃ = Δ + (σ²-Σ²)/2 ∇²Δ with ∇²Δ = (r²/Σ⁴ - 3/Σ²) Δ (test_spread_forces.jl:29-43)."""
    D = (2 * np.pi * Sigma**2) ** -1.5 * np.exp(-r2 / (2 * Sigma**2))
    return D + (sigma**2 - Sigma**2) / 2 * (r2 / Sigma**4 - 3 / Sigma**2) * D


@pytest.mark.parametrize("T", TYPES)
def test_spread_single_particle_closed_form(T):
    # test/accuracy/test_spread_forces.jl:7-52
    cfg = orc.OracleConfig(L=(4.0, 4.0, 4.0), R_c=1.0, N=1, kernel_widths_ratio=2.0, num_grid_points=(8, 8, 8),
                           M_G=8, viscosity=1.0, dtype=T)
    grid = orc.spread(cfg, [[2.0, 2.0, 2.0]], [[1.0, 0.0, 0.0]])
    h = float(cfg.h)
    g = np.arange(8) * h - 2.0
    r2 = g[None, None, :] ** 2 + g[None, :, None] ** 2 + g[:, None, None] ** 2
    expected = modified_kernel_f64(r2, float(cfg.sigma), float(cfg.Sigma))
    np.testing.assert_allclose(grid[0], expected, rtol=rtol_of(T))
    assert np.all(grid[1] == 0) and np.all(grid[2] == 0)


@pytest.mark.parametrize("T", TYPES)
def test_spread_anchor_known_answers(T):
    # test/accuracy/test_spread_forces.jl:54-93 (M = 16, M_G = 4)
    cfg = orc.OracleConfig(L=(8.0, 8.0, 8.0), R_c=1.0, N=1, kernel_widths_ratio=2.0, num_grid_points=(16, 16, 16),
                           M_G=4, viscosity=1.0, dtype=T)
    for frac, expected in ((0.3, [15, 16, 1, 2]), (0.7, [16, 1, 2, 3])):
        y = T(frac) * cfg.h
        grid = orc.spread(cfg, [[y, y, y]], [[1.0, 0.0, 0.0]])
        marginal_x = np.abs(grid[0]).sum(axis=(0, 1))
        assert sorted((np.nonzero(marginal_x > 0)[0] + 1).tolist()) == sorted(expected)


def test_spread_random_cloud_matches_closed_form_sum():
    # Each particle contributes Δ̃ at the grid points of its M_G^3 anchor-centred stencil.
    T = np.float64
    cfg = orc.OracleConfig(L=(6.0, 6.0, 6.0), R_c=1.5, N=5, kernel_widths_ratio=1.7, num_grid_points=(12, 12, 12),
                           M_G=6, viscosity=1.0, dtype=T)
    rng = np.random.default_rng(11)
    Y = rng.uniform(0, 6.0, size=(5, 3))
    F = rng.normal(size=(5, 3))
    grid = orc.spread(cfg, Y, F)
    expected = np.zeros_like(grid)
    h = float(cfg.h)
    for n in range(5):
        j = np.rint(Y[n] / h).astype(int)
        offs = np.arange(6) - 3
        for dz, dy, dx in itertools.product(offs, offs, offs):
            g = j + np.array([dx, dy, dz])
            x = g * h - Y[n]
            val = modified_kernel_f64(float(np.dot(x, x)), float(cfg.sigma), float(cfg.Sigma))
            iz, iy, ix = g[2] % 12, g[1] % 12, g[0] % 12
            expected[:, iz, iy, ix] += val * F[n]
    np.testing.assert_allclose(grid, expected, rtol=1e-12, atol=1e-14)


# ----------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("M", [(8, 8, 8), (7, 9, 5), (12, 10, 6), (15, 6, 25), (16, 1, 4), (1, 3, 2)])
@pytest.mark.parametrize("T", TYPES)
def test_fft_matches_numpy(M, T):
    rng = np.random.default_rng(5)
    Mx, My, Mz = M
    f = rng.normal(size=(Mz, My, Mx))
    spec = orc.rfft3(M, f.astype(T), T)
    ref = np.fft.rfftn(f.astype(T).astype(np.float64), axes=(0, 1, 2))
    tol = 2e-6 if T is np.float32 else 1e-13
    assert np.linalg.norm(spec - ref) / np.linalg.norm(ref) < tol
    # brfft of an arbitrary (non-Hermitian) half spectrum = M * numpy irfftn (Re of the full
    # inverse with the x DC/Nyquist imaginary parts dropped).
    X = rng.normal(size=(Mz, My, Mx // 2 + 1)) + 1j * rng.normal(size=(Mz, My, Mx // 2 + 1))
    out = orc.brfft3(M, X, T)
    ref = np.fft.irfftn(X.astype(np.complex64 if T is np.float32 else np.complex128).astype(np.complex128),
                        s=(Mz, My, Mx), axes=(0, 1, 2)) * (Mx * My * Mz)
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < tol


# ----------------------------------------------------------------------------- Stokes
@pytest.mark.parametrize("T", TYPES)
def test_stokes_single_mode_analytic(T):
    # test/accuracy/test_stokes_solve.jl:208-247: f = sin(k x) e_y -> u_y = sin(k x)/(μ k²)
    cfg = orc.OracleConfig(L=(4.0, 4.0, 4.0), R_c=1.0, N=1, kernel_widths_ratio=2.0, num_grid_points=(8, 8, 8),
                           M_G=8, viscosity=1.5, dtype=T)
    k = T(2) * T(np.pi) / cfg.L_T[0]
    x = np.arange(8, dtype=T) * cfg.h
    f = np.zeros((3, 8, 8, 8), dtype=T)
    f[1] = np.sin(k * x)[None, None, :]
    u, _ = orc.stokes(cfg, f)
    scale = 1 / (float(cfg.mu) * float(k) ** 2)
    atol = rtol_of(T) * scale
    assert np.all(np.abs(u[0]) <= atol) and np.all(np.abs(u[2]) <= atol)
    expected = np.broadcast_to(scale * np.sin(float(k) * x.astype(np.float64)), u[1].shape)
    np.testing.assert_allclose(u[1], expected, atol=atol)


@pytest.mark.parametrize("T", TYPES)
def test_stokes_divergence_free_and_zero(T):
    # test_stokes_solve.jl:150-206 (k·û = 0 per mode) and :515-537 (zero force -> exactly zero)
    cfg = orc.OracleConfig(L=(6.0, 5.0, 7.0), R_c=1.0, N=1, kernel_widths_ratio=2.0, num_grid_points=(12, 10, 14),
                           M_G=6, viscosity=0.8, dtype=T)
    rng = np.random.default_rng(2)
    f = rng.normal(size=(3, 14, 10, 12)).astype(T)
    u, fh = orc.stokes(cfg, f)
    k = np.stack(np.meshgrid(cfg.k_z, cfg.k_y, cfg.k_x, indexing="ij")[::-1]).astype(np.float64)
    div = (k * fh.astype(np.complex128)).sum(axis=0)
    assert np.max(np.abs(div)) <= 10 * rtol_of(T) * np.max(np.abs(fh)) * np.max(np.abs(k))
    assert abs(u.astype(np.float64).mean(axis=(1, 2, 3))).max() < 10 * rtol_of(T) * np.abs(u).max()
    u0, fh0 = orc.stokes(cfg, np.zeros_like(f))
    assert np.all(u0 == 0) and np.all(fh0 == 0)
    # The projected spectrum equals the projector applied to numpy's rfftn (float64).
    F = np.fft.rfftn(f.astype(np.float64), axes=(1, 2, 3))
    k2 = (k**2).sum(axis=0)
    k2s = np.where(k2 == 0, 1, k2)
    proj = (F - k * (k * F).sum(axis=0) / k2s) / (float(cfg.mu) * k2s) / (12 * 10 * 14)
    proj[:, k2 == 0] = 0
    assert np.linalg.norm(fh - proj) / np.linalg.norm(proj) < (5e-6 if T is np.float32 else 1e-12)


# ----------------------------------------------------------------------------- interpolation
def test_interpolation_is_adjoint_of_spread():
    # test/accuracy/test_interpolate_velocities.jl:8-63: <F, J u> = h³ <S F, u>
    T = np.float64
    cfg = orc.OracleConfig(L=(8.0, 8.0, 8.0), R_c=1.0, N=7, kernel_widths_ratio=2.0, num_grid_points=(16, 16, 16),
                           M_G=8, viscosity=1.0, dtype=T)
    rng = np.random.default_rng(4)
    Ys = rng.uniform(0, 8, size=(7, 3))
    F = rng.normal(size=(7, 3))
    u = rng.normal(size=(3, 16, 16, 16))
    orig = np.arange(1, 8, dtype=np.int32)
    V = orc.interpolate(cfg, u, Ys, orig)
    S = orc.spread(cfg, Ys, F)
    lhs = float((F * V).sum())
    rhs = float(cfg.h) ** 3 * float((S * u).sum())
    assert lhs == pytest.approx(rhs, rel=1e-12)
    # Original order: permuting the slots permutes the output rows.
    perm = np.array([3, 1, 7, 2, 5, 4, 6], dtype=np.int32)
    V2 = orc.interpolate(cfg, u, Ys, perm)
    np.testing.assert_array_equal(V2[perm - 1], V)


@pytest.mark.parametrize("T", TYPES)
def test_interpolation_constant_flow(T):
    # test/accuracy/test_interpolate_velocities.jl:185-228: a uniform flow interpolates to
    # itself (unit kernel mass) within sqrt(eps(T)) at L = 16, M = 16, M_G = 16, Σ/σ = 2.
    cfg = orc.OracleConfig(L=(16.0, 16.0, 16.0), R_c=1.0, N=4, kernel_widths_ratio=2.0,
                           num_grid_points=(16, 16, 16), M_G=16, viscosity=1.0, dtype=T)
    U = np.array([0.7, -1.3, 2.1], dtype=T)
    u = np.empty((3, 16, 16, 16), dtype=T)
    for c in range(3):
        u[c] = U[c]
    Ys = np.array([[3.3, 4.1, 8.0], [8.0, 8.0, 8.0], [11.7, 2.9, 13.4], [0.4, 15.6, 7.2]], dtype=T)
    V = orc.interpolate(cfg, u, Ys, np.arange(1, 5, dtype=np.int32))
    np.testing.assert_allclose(V, np.tile(U.astype(np.float64), (4, 1)), rtol=rtol_of(T), atol=atol_of(T))


# ----------------------------------------------------------------------------- correction
def _gaussian(r2, s2):
    return (2 * np.pi * s2) ** -1.5 * np.exp(-r2 / (2 * s2))


def _S(x, s, mu):
    r = np.linalg.norm(x)
    r2, s2 = r * r, s * s
    e = erf(r / (s * np.sqrt(2)))
    D = _gaussian(r2, s2)
    cI = e / (8 * np.pi * mu * r) * (1 + s2 / r2) - s2**2 / (2 * mu * r2) * D
    cxx = e / (8 * np.pi * mu * r**3) * (1 - 3 * s2 / r2) + 3 * s2**2 / (2 * mu * r**4) * D
    return cI * np.eye(3) + cxx * np.outer(x, x)


def _Q(x, s, mu):
    r = np.linalg.norm(x)
    r2, s2 = r * r, s * s
    e = erf(r / (s * np.sqrt(2)))
    D = _gaussian(r2, s2)
    cI = e / (4 * np.pi * mu * r**3) - (1 + s2 / r2) * D / mu
    cxx = -3 * e / (4 * np.pi * mu * r**5) + (1 + 3 * s2 / r2) / (mu * r2) * D
    return cI * np.eye(3) + cxx * np.outer(x, x)


def _Tt(x, s, mu):
    r = np.linalg.norm(x)
    r2, s2 = r * r, s * s
    D = _gaussian(r2, s2)
    return (2 - r2 / s2) / (mu * s2) * D * np.eye(3) + 1 / (mu * s2**2) * D * np.outer(x, x)


def correction_tensor(x, sigma, Sigma, mu):
    """test/accuracy/test_correct_velocities.jl:21-68 (independent full-tensor route)."""
    d = sigma**2 - Sigma**2
    ss, sS = sigma * np.sqrt(2), Sigma * np.sqrt(2)
    return _S(x, ss, mu) - _S(x, sS, mu) - d * _Q(x, sS, mu) - d**2 / 4 * _Tt(x, sS, mu)


@pytest.mark.parametrize("T", TYPES)
def test_correction_scalars_vs_full_tensor(T):
    sigma = 1 / np.sqrt(np.pi)
    dirs = [np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.ones(3) / np.sqrt(3),
            np.array([2, -1, 0.5]) / np.linalg.norm([2, -1, 0.5])]
    for ratio, mu in itertools.product((1.5, 2.5), (0.7, 1.6)):
        Sigma = ratio * sigma
        for r, d in itertools.product((0.3, 0.6, 1.2), dirs):
            A, B = orc.correction_scalars(r, sigma, Sigma, mu, T)
            x = r * d
            impl = float(A) * np.eye(3) + float(B) * np.outer(x, x)
            ref = correction_tensor(x, sigma, Sigma, mu)
            np.testing.assert_allclose(impl, ref, rtol=rtol_of(T), atol=atol_of(T))
    # Σ = σ: both scalars and the self term vanish (test_correct_velocities.jl:70-86)
    s = T(1) / np.sqrt(T(np.pi))
    for mu in (0.5, 1.0, 2.7):
        assert abs(orc.self_correction(s, s, T(1), mu, T)) <= atol_of(T)
        for r in (0.2, 0.75, 1.5):
            A, B = orc.correction_scalars(r, s, s, mu, T)
            assert abs(A) <= atol_of(T) and abs(B) <= atol_of(T)


@pytest.mark.parametrize("T", TYPES)
def test_self_correction_grouped_form(T):
    # test_correct_velocities.jl:114-137 (paper eq. B.1 grouped term by term)
    sigma = 1 / np.sqrt(np.pi)
    for ratio, mu in itertools.product((1.25, 2.0, 3.5), (0.5, 1.0, 2.7)):
        S = ratio * sigma
        Ssp = S * np.sqrt(np.pi)
        d = sigma**2 - S**2
        ref = 1 / (6 * np.pi * mu) - 1 / (6 * np.pi * mu * Ssp) + d / (12 * mu * Ssp**3) - d**2 / (32 * mu * S**5 * np.pi**1.5)
        got = orc.self_correction(sigma, S, 1.0, mu, T)
        assert float(got) == pytest.approx(ref, rel=rtol_of(T))


def _apply_correction(cfg, Y, F):
    Yw = orc.wrap_positions(cfg, Y)
    h = orc.assign_cells(cfg, Yw)
    orig, cs, ce = orc.build_cell_list(h, cfg.total_cells)
    Ys, Fs = orc.gather(cfg, orig, Yw, F)
    V = np.zeros((cfg.N, 3), dtype=cfg.T)
    return orc.correct(cfg, Ys, Fs, cs, ce, orig, V)


def _std_cfg(T, N, **kw):
    """test/test_utilities.jl:35-53 (_standard_test_config)."""
    base = dict(L=(8.0, 8.0, 8.0), R_c=1.0, N=N, kernel_widths_ratio=2.0, num_grid_points=(16, 16, 16), M_G=8,
                viscosity=1.0)
    base.update(kw)
    return orc.OracleConfig(dtype=T, **base)


@pytest.mark.parametrize("T", TYPES)
def test_correction_isolated_and_pair(T):
    # test_correct_velocities.jl:139-173
    cfg = _std_cfg(T, 2)
    c = float(orc.self_correction(cfg.sigma, cfg.Sigma, cfg.a_T, cfg.mu, T))
    F = np.array([[0.3, -0.7, 1.1], [-0.5, 0.9, 0.2]])
    V = _apply_correction(cfg, np.array([[1, 1, 1], [4, 4, 4]], dtype=T), F)
    np.testing.assert_allclose(V, c * F, rtol=rtol_of(T), atol=atol_of(T))
    Y = np.array([[4, 4, 4], [4.5, 4.2, 3.8]], dtype=T)
    V = _apply_correction(cfg, Y, F)
    x = (Y[0] - Y[1]).astype(np.float64)
    Mt = correction_tensor(x, float(cfg.sigma), float(cfg.Sigma), float(cfg.mu))
    np.testing.assert_allclose(V[0], c * F[0] + Mt @ F[1], rtol=rtol_of(T), atol=atol_of(T))
    np.testing.assert_allclose(V[1], c * F[1] + Mt @ F[0], rtol=rtol_of(T), atol=atol_of(T))


def test_correction_symmetry_translation_image():
    # test_correct_velocities.jl:175-226
    T = np.float64
    cfg = _std_cfg(T, 5)
    Y = np.array([[4.0, 4.0, 4.0], [4.4, 3.8, 4.2], [3.7, 4.3, 4.0], [4.2, 4.1, 3.7], [3.9, 3.6, 4.4]])
    Fa = np.array([[0.3, -0.7, 1.1], [-0.5, 0.9, 0.2], [0.8, -0.1, -0.6], [-0.2, 0.5, 0.3], [0.6, -0.4, 0.7]])
    Fb = np.array([[-0.2, 0.5, -0.4], [0.6, -0.3, 0.1], [-0.9, 0.7, 0.3], [0.4, -0.8, 0.6], [0.1, 0.2, -0.5]])
    Va = _apply_correction(cfg, Y, Fb)
    Vb = _apply_correction(cfg, Y, Fa)
    assert (Fa * Va).sum() == pytest.approx((Fb * Vb).sum(), rel=1e-12)
    V_ref = _apply_correction(cfg, Y, Fa)
    np.testing.assert_allclose(_apply_correction(cfg, Y + np.array([1.3, -2.1, 0.7]), Fa), V_ref, rtol=1e-8, atol=1e-10)
    Yi = Y.copy()
    Yi[1, 0] += 8.0
    np.testing.assert_allclose(_apply_correction(cfg, Yi, Fa), V_ref, rtol=1e-8, atol=1e-10)


# ----------------------------------------------------------------------------- end to end
def periodic_self_mobility_xx(sigma, mu, L, n_max=15):
    """test/test_utilities.jl:112-123 (reciprocal-lattice sum)."""
    n = np.arange(-n_max, n_max + 1)
    nx, ny, nz = np.meshgrid(n, n, n, indexing="ij")
    k = 2 * np.pi / L * np.stack([nx, ny, nz]).astype(np.float64)
    k2 = (k**2).sum(axis=0)
    mask = k2 > 0
    terms = np.exp(-sigma**2 * k2[mask]) / (mu * k2[mask]) * (1 - k[0][mask] ** 2 / k2[mask])
    return terms.sum() / L**3


def test_single_sphere_self_mobility():
    # test/accuracy/test_single_sphere_mobility.jl:6-64 (Σ = σ, M = 32, M_G = 21, rtol 1e-5)
    T = np.float64
    cfg = orc.OracleConfig(L=(8.0, 8.0, 8.0), R_c=1.0, N=1, kernel_widths_ratio=1.0, num_grid_points=(32, 32, 32),
                           M_G=21, viscosity=1.0, dtype=T)
    V = orc.mobility(cfg, [[4.0, 4.0, 4.0]], [[1.0, 0.0, 0.0]])
    ref = periodic_self_mobility_xx(float(cfg.sigma), 1.0, 8.0)
    assert V[0, 0] == pytest.approx(ref, rel=1e-5)
    assert abs(V[0, 1]) < 1e-12 and abs(V[0, 2]) < 1e-12
    assert 0 < V[0, 0] < 1 / (6 * np.pi)


def test_single_sphere_sigma_independent():
    # test_single_sphere_mobility.jl:66-110 (Σ/σ ∈ {1.25, 1.5}, M_G = 31, rtol 1e-4)
    T = np.float64
    sigma = 1 / np.sqrt(np.pi)
    ref = periodic_self_mobility_xx(sigma, 1.0, 8.0)
    vs = []
    for ratio in (1.25, 1.5):
        cfg = orc.OracleConfig(L=(8.0, 8.0, 8.0), R_c=1.0, N=1, kernel_widths_ratio=ratio,
                               num_grid_points=(32, 32, 32), M_G=31, viscosity=1.0, dtype=T)
        V = orc.mobility(cfg, [[4.0, 4.0, 4.0]], [[1.0, 0.0, 0.0]])
        assert V[0, 0] == pytest.approx(ref, rel=1e-4)
        vs.append(V[0, 0])
    assert vs[0] == pytest.approx(vs[1], rel=1e-4)


@pytest.mark.parametrize("T", TYPES)
def test_mobility_operator_identities(T):
    # test/accuracy/test_mobility_properties.jl:39-69 and test_mobility.jl:6-145 (own seeds)
    rng = np.random.default_rng(2024)
    for _ in range(3):
        L = float(rng.uniform(6, 10))
        M = int(rng.integers(12, 20))
        N = int(rng.integers(2, 12))
        cfg = orc.OracleConfig(L=(L, L, L), R_c=min(1.5, L / 2), N=N, kernel_widths_ratio=float(rng.uniform(1, 2.5)),
                               num_grid_points=(M, M, M), M_G=int(rng.integers(6, min(M, 10) + 1)),
                               viscosity=float(rng.uniform(0.5, 2)), dtype=T)
        Y = ((3 * rng.random((N, 3)) - 1) * L).astype(T)
        F = (2 * rng.random((N, 3)) - 1).astype(T)
        G = (2 * rng.random((N, 3)) - 1).astype(T)
        a, b = T(rng.uniform(-1, 1)), T(rng.uniform(-1, 1))
        VF = orc.mobility(cfg, Y, F)
        VG = orc.mobility(cfg, Y, G)
        Vc = orc.mobility(cfg, Y, a * F + b * G)
        np.testing.assert_allclose(Vc, a * VF + b * VG, rtol=rtol_of(T), atol=atol_of(T))
        assert float((G * VF).sum()) == pytest.approx(float((F * VG).sum()), rel=rtol_of(T), abs=atol_of(T))
        assert float((F * VF).sum()) > 0
        Y0 = Y.copy()
        orc.mobility(cfg, Y, F)
        assert np.array_equal(Y, Y0)


def test_float32_tracks_float64():
    rng = np.random.default_rng(9)
    N, L = 64, 12.0
    Y = rng.uniform(0, L, size=(N, 3))
    F = rng.normal(size=(N, 3))
    F -= F.mean(axis=0)
    kw = dict(L=(L, L, L), R_c=3.0, N=N, kernel_widths_ratio=2.0, num_grid_points=(24, 24, 24), M_G=8, viscosity=1.0)
    V32 = orc.mobility(orc.OracleConfig(dtype=np.float32, **kw), Y.astype(np.float32), F.astype(np.float32))
    V64 = orc.mobility(orc.OracleConfig(dtype=np.float64, **kw), Y.astype(np.float32).astype(np.float64),
                       F.astype(np.float32).astype(np.float64))
    assert orc.rel_l2(V32, V64) < 1e-5


# ----------------------------------------------------------------------------- adjoint
# The pullback of mobility! has no reference function; its oracle (mobility_vjp) is pinned by
# float64 central differences of the reference-restated forward path, never by its own
# formulas.
def test_correction_scalar_derivatives_match_central_differences():
    T = np.float64
    sigma = 1 / np.sqrt(np.pi)
    for ratio in (1.25, 2.0, 3.0):
        Sigma = ratio * sigma
        for r in np.linspace(0.35, 4.5, 23):
            eps = 1e-5
            Ap, Bp = orc.correction_scalars(r + eps, sigma, Sigma, 1.3, T)
            Am, Bm = orc.correction_scalars(r - eps, sigma, Sigma, 1.3, T)
            dA, dB = orc.correction_scalars_deriv(r, sigma, Sigma, 1.3, T)
            assert float(dA) == pytest.approx((Ap - Am) / (2 * eps), rel=1e-6, abs=1e-12)
            assert float(dB) == pytest.approx((Bp - Bm) / (2 * eps), rel=1e-6, abs=1e-12)


def _rsa(rng, N, L, dmin):
    Y = []
    while len(Y) < N:
        y = rng.uniform(0, L, 3)
        ok = True
        for q in Y:
            d = y - q
            d -= L * np.round(d / L)
            if d @ d < dmin * dmin:
                ok = False
                break
        if ok:
            Y.append(y)
    return np.array(Y)


def _adjoint_system(seed, N=20, L=12.0, M=24, R_c=3.0, ratio=2.0):
    rng = np.random.default_rng(seed)
    cfg = orc.OracleConfig(L=(L, L, L), R_c=R_c, N=N, kernel_widths_ratio=ratio, num_grid_points=(M, M, M), M_G=8,
                           viscosity=1.0, dtype=np.float64)
    Y = _rsa(rng, N, L, 2.0)
    F = rng.normal(size=(N, 3))
    G = rng.normal(size=(N, 3))
    return cfg, Y, F, G


def _smooth_at(cfg, Y, n, i, eps):
    """True when Y[n, i] +- eps moves no stencil anchor and no pair of n across R_c."""
    h = float(cfg.h)
    for s in (-eps, eps):
        if np.rint((Y[n, i] + s) / h) != np.rint(Y[n, i] / h):
            return False
    L = float(cfg.L_T[0])
    Rc = float(cfg.R_c_T)
    for m in range(len(Y)):
        if m == n:
            continue
        r = []
        for s in (-eps, 0.0, eps):
            d = Y[n] - Y[m]
            d[i] += s
            d -= L * np.round(d / L)
            r.append(np.sqrt(d @ d))
        if (min(r) - Rc) * (max(r) - Rc) <= 0:
            return False
    return True


def test_vjp_dF_is_mobility_of_G_and_symmetric():
    cfg, Y, F, G = _adjoint_system(11)
    dF, _ = orc.mobility_vjp(cfg, Y, F, G)
    np.testing.assert_array_equal(dF, orc.mobility(cfg, Y, G))
    # dL/dF = M^T G; M^T = M is the operator identity <G, M F> = <F, M G>
    assert float((G * orc.mobility(cfg, Y, F)).sum()) == pytest.approx(float((F * dF).sum()), rel=1e-12)


@pytest.mark.parametrize("seed,ratio", [(3, 2.0), (4, 1.5)])
def test_vjp_dY_matches_central_differences(seed, ratio):
    cfg, Y, F, G = _adjoint_system(seed, ratio=ratio)
    _, dY = orc.mobility_vjp(cfg, Y, F, G)
    eps = 1e-4
    scale = float(np.sqrt((dY**2).mean()))
    checked = 0
    for n in range(cfg.N):
        for i in range(3):
            if not _smooth_at(cfg, Y, n, i, eps):
                continue
            Yp = Y.copy()
            Ym = Y.copy()
            Yp[n, i] += eps
            Ym[n, i] -= eps
            fd = ((G * orc.mobility(cfg, Yp, F)).sum() - (G * orc.mobility(cfg, Ym, F)).sum()) / (2 * eps)
            assert dY[n, i] == pytest.approx(fd, rel=1e-5, abs=1e-6 * scale), (n, i)
            checked += 1
    assert checked >= 30


def test_vjp_dY_pair_and_grid_parts_both_present():
    # A pair inside R_c and an isolated particle: the isolated one sees only the grid part, so a
    # pair-part error and a grid-part error cannot hide behind each other in one sum.
    cfg, Y, F, G = _adjoint_system(5, N=3)
    Y = np.array([[3.0, 3.1, 2.9], [4.6, 3.4, 3.3], [9.0, 8.7, 9.2]])
    _, dY = orc.mobility_vjp(cfg, Y, F, G)
    eps = 1e-4
    for n in range(3):
        for i in range(3):
            assert _smooth_at(cfg, Y, n, i, eps)
            Yp = Y.copy()
            Ym = Y.copy()
            Yp[n, i] += eps
            Ym[n, i] -= eps
            fd = ((G * orc.mobility(cfg, Yp, F)).sum() - (G * orc.mobility(cfg, Ym, F)).sum()) / (2 * eps)
            assert dY[n, i] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_calibration_selection_and_candidates():
    # the calibration's choice logic (ForceCouplingMethod/calibration.py) with a synthetic error
    # and cost: the cheapest candidate within tolerance wins, a tighter tolerance never costs
    # less, grids are 5-smooth, infeasible sets are refused
    from ForceCouplingMethod import calibration as cal
    from ForceCouplingMethod.suspensions import smooth5_at_least

    c = cal.Candidate(2.0, 1.5, 8, 3.0)
    M = c.grid(140.0)
    assert M == smooth5_at_least(M) and 140.0 / M <= c.Sigma / 1.5 + 1e-12
    assert cal.Candidate(2.0, 1.0, 8, 4.0).feasible(140.0)
    assert not cal.Candidate(3.0, 2.0, 12, 4.0).feasible(10.0)  # R_c above L/2
    cands = [cal.Candidate(r, sh, mg, cut) for r in cal.RATIOS for sh in cal.SIGMA_PER_H for mg in cal.STENCILS
             for cut in cal.CUTOFFS]
    err = lambda c: 10.0 ** (-c.sigma_per_h * c.M_G / 4.0) + 10.0 ** (-c.cutoff - 1.5)
    cost = lambda c: c.grid(100.0) ** 3 * 1e-6 + c.M_G ** 3 * 1e-3 + (c.cutoff * c.Sigma) ** 3 * 1e-2
    prev = 0.0
    for tol in (1e-2, 1e-3, 1e-4):
        best = cal.select(cands, err, cost, tol)
        assert best is not None
        choice, t, e = best
        assert e <= tol and t >= prev
        assert all(cost(x) >= t for x in cands if err(x) <= tol)
        prev = t
    assert cal.select(cands, err, cost, 1e-30) is None
    m = cal.CostModel({"cell": [1e-6, 0.01], "spread": [1e-9, 1e-9, 0.0], "fft": [1e-10, 0.0],
                       "interp": [1e-9, 1e-9, 0.0], "correct": [1e-7, 1e-7, 0.0]})
    assert m.predict(65536, 140.0, 256, 8, 4.0) < m.predict(65536, 140.0, 512, 8, 4.0)
