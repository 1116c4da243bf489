"""Per-step CPU <-> GPU parity, mirroring test/cuda/test_gpu_*.jl: identical inputs are written
into the GPU config's buffers and into the oracle, the step runs on both, and the results are
compared. Integer results (wrap, hashes, permutation, cell ranges, gathered rows) must be
bit-identical; floating-point grids must agree to relative L2 <= sqrt(eps(T)) (the
reference's isapprox norm test, test/test_utilities.jl:65-67)."""
import numpy as np
import pytest

from helpers import kwargs, oracle_grid, oracle_rows, pair, rel_l2, rtol_of
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TYPES = [np.float32, np.float64]


def _random_cloud(rng, N, L, T, lo=-1.0, hi=2.0):
    Y = np.empty((3, N), dtype=T, order="F")
    for i in range(3):
        Y[i] = rng.uniform(lo * L[i], hi * L[i], size=N).astype(T)
    F = np.asfortranarray(rng.normal(size=(3, N)).astype(T))
    return Y, F


# ---------------------------------------------------------------- (a) wrap, hash, cell list
@pytest.mark.parametrize("T", TYPES)
def test_wrap_and_hash_bit_exact(T):
    from ForceCouplingMethod import assign_cells_, wrap_positions_

    L = (4.0, 6.0, 8.0)
    N = 4000
    gpu, ora = pair(T, L=L, R_c=1.0, N=N, M=(8, 12, 16))
    rng = np.random.default_rng(11)
    Y, _ = _random_cloud(rng, N, L, T)
    # the edge cases of test/accuracy/test_wrap_positions.jl:9-62
    eps = np.finfo(T).eps
    edge = [0.0, L[0], -L[0], 2 * L[0], -float(eps), L[0] * (1 - eps), 3 * L[0] + 0.5, -0.0]
    for j, v in enumerate(edge):
        Y[0, j] = T(v)
        Y[1, j] = T(v * L[1] / L[0])
        Y[2, j] = T(v * L[2] / L[0])
    gpu.set_buffer("Y_input", Y)
    wrap_positions_(gpu)
    Yw = gpu.get_buffer("Y_wrapped")
    Yw_ref = orc.wrap_positions(ora, oracle_rows(Y)).T
    assert np.array_equal(Yw, Yw_ref)
    assert np.array_equal(np.signbit(Yw), np.signbit(Yw_ref))
    assign_cells_(gpu)
    h = gpu.get_buffer("cell_hash")
    assert np.array_equal(h, orc.assign_cells(ora, oracle_rows(Yw_ref)))


def test_hash_known_answers():
    # test/accuracy/test_assign_cells_kernel.jl:5-63
    from ForceCouplingMethod import assign_cells_

    T = np.float32
    cases = [((4.0, 4.0, 4.0), [(0.5, 0.5, 0.5), (2.5, 2.5, 2.5), (3.999, 3.999, 3.999), (4.0, 4.0, 4.0)], [0, 42, 63, 63]),
             ((4.0, 6.0, 8.0), [(0.5, 2.5, 5.5)], [128])]
    for L, pts, want in cases:
        M = tuple(int(round(l / (L[0] / 8))) for l in L)
        gpu, _ = pair(T, L=L, R_c=1.0, N=len(pts), M=M)
        gpu.set_buffer("Y_wrapped", np.asfortranarray(np.array(pts, dtype=T).T))
        assign_cells_(gpu)
        assert list(gpu.get_buffer("cell_hash")) == want


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("case", ["uniform", "clustered", "single_cell", "many_cells", "mid_counts", "million"])
def test_cell_list_bit_exact(T, case):
    from ForceCouplingMethod import assign_cells_, sort_particles_by_cell_, wrap_positions_

    rng = np.random.default_rng(hash(case) % 2**32)
    if case == "uniform":
        L, Rc, N, M = (8.0, 8.0, 8.0), 1.0, 9000, (16, 16, 16)
    elif case == "clustered":
        L, Rc, N, M = (8.0, 8.0, 8.0), 1.0, 64, (16, 16, 16)
    elif case == "single_cell":
        L, Rc, N, M = (6.0, 6.0, 6.0), 2.0, 5000, (12, 12, 12)
    elif case == "many_cells":  # 2^17 cells: three radix passes
        L, Rc, N, M = (64.0, 64.0, 128.0), 1.0, 70000, (64, 64, 128)
    elif case == "mid_counts":  # 98 sort tiles: 25,088 digit counts, scanned by one block (above 2^14)
        L, Rc, N, M = (64.0, 64.0, 64.0), 1.0, 200000, (64, 64, 64)
    else:  # 2^20 + 37 particles: the sort's 4096-key tiles, the last one partial
        L, Rc, N, M = (64.0, 64.0, 64.0), 1.0, (1 << 20) + 37, (64, 64, 64)
    gpu, ora = pair(T, L=L, R_c=Rc, N=N, M=M)
    if case == "clustered":
        n = np.arange(N)  # _clustered_cloud (test/test_utilities.jl:93-105)
        Y = np.stack([4 + 0.25 * (n % 4), 4 + 0.25 * ((n // 4) % 4), 4 + 0.25 * (n // 16)]).astype(T)
        F = np.stack([0.1 * (n + 1), -0.2 * (n + 1), 0.3 * (n + 1)]).astype(T)
    elif case == "single_cell":
        Y = rng.uniform(0.1, 1.9, size=(3, N)).astype(T)
        F = rng.normal(size=(3, N)).astype(T)
    else:
        Y, F = _random_cloud(rng, N, L, T)
    Y, F = np.asfortranarray(Y), np.asfortranarray(F)
    gpu.set_buffer("Y_input", Y)
    gpu.set_buffer("F_input", F)
    wrap_positions_(gpu)
    assign_cells_(gpu)
    sort_particles_by_cell_(gpu)
    Yw = orc.wrap_positions(ora, oracle_rows(Y))
    h = orc.assign_cells(ora, Yw)
    orig, cs, ce = orc.build_cell_list(h, ora.total_cells)
    Ys, Fs = orc.gather(ora, orig, Yw, oracle_rows(F))
    assert np.array_equal(gpu.get_buffer("cell_hash"), h)
    assert np.array_equal(gpu.get_buffer("original_index"), orig)
    assert np.array_equal(gpu.get_buffer("cell_start"), cs)
    assert np.array_equal(gpu.get_buffer("cell_end"), ce)
    assert np.array_equal(gpu.get_buffer("Y_sorted"), Ys.T)
    assert np.array_equal(gpu.get_buffer("F_sorted"), Fs.T)


def test_sort_known_answers():
    # test/accuracy/test_sort_particles_by_cell.jl:7-72: hashes [0,1,0,1,0] -> cell 0 = [1,3,5]
    from ForceCouplingMethod import sort_particles_by_cell_

    T = np.float32
    gpu, _ = pair(T, L=(4.0, 4.0, 4.0), R_c=1.0, N=5, M=(8, 8, 8))
    Y = np.asfortranarray(np.array([[0.5, 0.5, 0.5], [1.5, 0.5, 0.5], [0.5, 0.5, 0.5], [1.5, 0.5, 0.5],
                                    [0.5, 0.5, 0.5]], dtype=T).T)
    gpu.set_buffer("Y_wrapped", Y)
    gpu.set_buffer("cell_hash", np.array([0, 1, 0, 1, 0], dtype=np.int32))
    gpu.set_buffer("F_input", np.zeros((3, 5), T))
    sort_particles_by_cell_(gpu)
    assert list(gpu.get_buffer("original_index")) == [1, 3, 5, 2, 4]
    cs, ce = gpu.get_buffer("cell_start"), gpu.get_buffer("cell_end")
    assert (cs[0], ce[0], cs[1], ce[1]) == (1, 3, 4, 5)
    assert cs[2] == 6 and ce[2] == 5  # empty cell: end = start - 1


def test_neighbor_map_matches_host_build():
    # test/cuda/test_gpu_cell_list.jl:96-98
    T = np.float32
    gpu, _ = pair(T, L=(5.0, 6.0, 7.0), R_c=1.0, N=3, M=(10, 12, 14))
    nm = gpu.get_buffer("neighbor_map").reshape(-1, 13)
    mx, my, mz = gpu.num_cells
    c = 0 + (1 + 2 * my) * mx  # cell (0, 1, 2)
    assert nm[c][0] == 1 + (1 + 2 * my) * mx  # (+1, 0, 0)
    assert nm[c][3] == (mx - 1) + (2 + 2 * my) * mx  # (-1, +1, 0) wraps in x


# ---------------------------------------------------------------- (b) spreading
def _sorted_inputs(ora, Y, F):
    Yw = orc.wrap_positions(ora, oracle_rows(Y))
    h = orc.assign_cells(ora, Yw)
    orig, cs, ce = orc.build_cell_list(h, ora.total_cells)
    Ys, Fs = orc.gather(ora, orig, Yw, oracle_rows(F))
    return Ys, Fs, orig, cs, ce


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("geom", [
    dict(L=(8.0, 8.0, 8.0), M=(16, 16, 16), M_G=8, N=300),       # one tile
    dict(L=(32.0, 32.0, 32.0), M=(64, 64, 64), M_G=8, N=3000),   # 4x4x8 tiles of 16x16x8
    dict(L=(20.0, 20.0, 20.0), M=(40, 40, 40), M_G=8, N=2000),   # 2x2x2 tiles of 20
    dict(L=(12.0, 24.0, 18.0), M=(24, 48, 36), M_G=6, N=1500),   # generic M_G, mixed tiles
    dict(L=(24.0, 24.0, 24.0), M=(48, 48, 48), M_G=10, N=2500),  # generic M_G, two lane steps
    dict(L=(16.0, 16.0, 8.0), M=(64, 64, 32), M_G=16, N=600),    # generic M_G, four lane steps
    dict(L=(8.0, 8.0, 8.0), M=(32, 32, 32), M_G=21, N=40),       # wide stencil
    dict(L=(16.0, 16.0, 16.0), M=(32, 32, 32), M_G=8, N=3000),   # 2 tiles per axis, many batches
    dict(L=(32.0, 16.0, 12.0), M=(64, 32, 24), M_G=8, N=2500),   # mixed extents, 3 planes of tiles
    dict(L=(16.0, 16.0, 4.0), M=(32, 32, 8), M_G=8, N=600),      # one tile plane: 4 home tiles, z wraps
    dict(L=(30.0, 18.0, 12.0), M=(60, 36, 24), M_G=8, N=2000),   # 15 x 12 x 8 tiles (partial products)
    dict(L=(16.0, 64.0, 16.0), M=(32, 128, 32), M_G=8, N=4000),  # 8 tile rows: per-XCD tile queues
    dict(L=(24.0, 128.0, 8.0), M=(48, 256, 16), M_G=8, N=3000),  # 3 x 16 x 2 tiles: 16 rows, 2 layers
])
@pytest.mark.parametrize("mode", ["default", "lds", "queue"])
def test_spread_matches_oracle(T, geom, mode, monkeypatch):
    # Float32 with M_G = 8 on Tx x Ty x 8 tiles (8 <= Tx, Ty <= 16, grids of at least T + 16 points
    # along x and y) spreads on the matrix cores by default, in a static tile order at these sizes;
    # FFCM_SPREAD_ORDER=queue draws the tiles from per-XCD queues where the tile rows are a
    # multiple of 8 (the order of large grids), and FFCM_SPREAD=lds selects the LDS accumulation
    # for the same tiles
    from ForceCouplingMethod import spread_forces_

    monkeypatch.delenv("FFCM_SPREAD_ORDER", raising=False)
    if mode == "lds":
        monkeypatch.setenv("FFCM_SPREAD", "lds")
    else:
        monkeypatch.delenv("FFCM_SPREAD", raising=False)
    if mode == "queue":
        monkeypatch.setenv("FFCM_SPREAD_ORDER", "queue")
    rng = np.random.default_rng(5)
    gpu, ora = pair(T, L=geom["L"], R_c=1.0, N=geom["N"], M=geom["M"], M_G=geom["M_G"])
    Y, F = _random_cloud(rng, geom["N"], geom["L"], T, 0.0, 1.0)
    Ys, Fs, orig, cs, ce = _sorted_inputs(ora, Y, F)
    gpu.set_buffer("Y_sorted", Ys.T)
    gpu.set_buffer("F_sorted", Fs.T)
    spread_forces_(gpu)
    got = gpu.get_buffer("force_density")
    ref = oracle_grid(orc.spread(ora, Ys, Fs))
    assert rel_l2(got, ref) <= rtol_of(T) * (1e-3 if T is np.float64 else 0.05)


def test_spread_anchor_known_answers():
    # test/accuracy/test_spread_forces.jl:54-93: M = 16, M_G = 4, Y/h = 0.3 -> x-support {15,16,1,2},
    # Y/h = 0.7 -> {16,1,2,3} (1-based)
    from ForceCouplingMethod import spread_forces_

    T = np.float64
    for frac, support in ((0.3, {15, 16, 1, 2}), (0.7, {16, 1, 2, 3})):
        gpu, _ = pair(T, L=(8.0, 8.0, 8.0), R_c=1.0, N=1, M=(16, 16, 16), M_G=4)
        h = 0.5
        gpu.set_buffer("Y_sorted", np.array([[frac * h], [4.0], [4.0]], dtype=T))
        gpu.set_buffer("F_sorted", np.array([[1.0], [0.0], [0.0]], dtype=T))
        spread_forces_(gpu)
        fx = gpu.get_buffer("force_density")[0]
        nz = set((np.nonzero(np.abs(fx).sum(axis=(1, 2)))[0] + 1).tolist())
        assert nz == support


# ---------------------------------------------------------------- (c) FFT Stokes solve
# (48, 90, 60, 120, 81, 75, 50, 144, 150, 160: compile-time Float32 radix plans with radix-3/5
# stages; 7, 26, 30, 12, 24, 36, 40: the runtime-radix plan)
GRIDS = [(16, 16, 16), (64, 64, 64), (7, 7, 7), (30, 30, 30), (12, 24, 36), (40, 40, 40), (26, 26, 26),
         (48, 48, 48), (90, 60, 120), (81, 75, 50), (144, 150, 160)]


def _grid_cfg(T, M):
    h = 0.5
    return pair(T, L=tuple(h * m for m in M), R_c=1.0, N=1, M=M, M_G=min(6, min(M)), mu=1.3)


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("M", GRIDS)
def test_stokes_forward_projection_matches_oracle(T, M):
    from ForceCouplingMethod import stokes_forward_project_, stokes_solve_

    gpu, ora = _grid_cfg(T, M)
    rng = np.random.default_rng(sum(M))
    f = rng.normal(size=(3,) + M).astype(T)
    gpu.set_buffer("force_density", f)
    stokes_forward_project_(gpu)
    fh = gpu.get_buffer("fluid_hat")
    u_ref, fh_ref = orc.stokes(ora, np.ascontiguousarray(np.transpose(f, (0, 3, 2, 1))))
    fh_ref = np.transpose(fh_ref, (0, 3, 2, 1))
    tol = 2e-6 if T is np.float32 else 1e-13
    assert rel_l2(fh, fh_ref) <= tol
    gpu.set_buffer("force_density", f)
    stokes_solve_(gpu)
    u = gpu.get_buffer("fluid_velocity")
    assert rel_l2(u, oracle_grid(u_ref)) <= tol


# every length with a compile-time Float32 plan (csrc/fft_plans_*.hip)
STATIC_LENGTHS = [48, 50, 54, 60, 64, 72, 75, 80, 81, 90, 96, 100, 108, 120, 125, 128, 135, 144, 150, 160, 162,
                  180, 192, 200, 216, 225, 240, 243, 250, 256, 270, 288, 300, 320, 324, 384, 512]


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("n", STATIC_LENGTHS)
def test_every_static_plan_on_each_axis(n, axis):
    # the x, y and z passes of each compile-time plan against the oracle's transform, with
    # small companion axes (16 and 18 points: their own plans are covered by the cubic grids)
    from ForceCouplingMethod import stokes_forward_project_, stokes_solve_

    T = np.float32
    M = [16, 18]
    M.insert(axis, n)
    M = tuple(M)
    gpu, ora = _grid_cfg(T, M)
    rng = np.random.default_rng(n * 3 + axis)
    f = rng.normal(size=(3,) + M).astype(T)
    gpu.set_buffer("force_density", f)
    stokes_forward_project_(gpu)
    fh = gpu.get_buffer("fluid_hat")
    u_ref, fh_ref = orc.stokes(ora, np.ascontiguousarray(np.transpose(f, (0, 3, 2, 1))))
    fh_ref = np.transpose(fh_ref, (0, 3, 2, 1))
    assert rel_l2(fh, fh_ref) <= 2e-6
    gpu.set_buffer("force_density", f)
    stokes_solve_(gpu)
    assert rel_l2(gpu.get_buffer("fluid_velocity"), oracle_grid(u_ref)) <= 2e-6


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("M", [(16, 16, 16), (7, 7, 7), (30, 30, 30), (12, 24, 36)])
def test_backward_c2r_matches_oracle_on_non_hermitian_input(T, M):
    # the c2r convention: complex passes along z and y, then the x rows as Hermitian with the
    # imaginary parts of the DC and Nyquist bins dropped
    from ForceCouplingMethod import stokes_backward_

    gpu, ora = _grid_cfg(T, M)
    rng = np.random.default_rng(3)
    C = np.complex64 if T is np.float32 else np.complex128
    spec = (rng.normal(size=(3, M[0] // 2 + 1, M[1], M[2])) + 1j * rng.normal(size=(3, M[0] // 2 + 1, M[1], M[2]))).astype(C)
    gpu.set_buffer("fluid_hat", spec)
    stokes_backward_(gpu)
    u = gpu.get_buffer("fluid_velocity")
    tol = 2e-6 if T is np.float32 else 1e-13
    for c in range(3):
        ref = orc.brfft3(M, np.ascontiguousarray(np.transpose(spec[c], (2, 1, 0))), T)
        assert rel_l2(u[c], np.transpose(ref, (2, 1, 0))) <= tol


def test_stokes_single_mode_analytic():
    # test/accuracy/test_stokes_solve.jl:208-247: f_y = sin(k x) -> u_y = sin(k x) / (mu k^2)
    from ForceCouplingMethod import stokes_solve_

    T = np.float64
    M = (16, 16, 16)
    gpu, _ = _grid_cfg(T, M)
    L = 8.0
    k = 2 * np.pi / L
    x = np.arange(16) * 0.5
    f = np.zeros((3,) + M)
    f[1] = np.sin(k * x)[:, None, None]
    gpu.set_buffer("force_density", f)
    stokes_solve_(gpu)
    u = gpu.get_buffer("fluid_velocity")
    np.testing.assert_allclose(u[1], f[1] / (1.3 * k * k), atol=1e-12)
    np.testing.assert_allclose(u[0], 0, atol=1e-12)
    np.testing.assert_allclose(u[2], 0, atol=1e-12)


def test_stokes_zero_force_is_exactly_zero():
    # test/accuracy/test_stokes_solve.jl:515-537
    from ForceCouplingMethod import stokes_solve_

    gpu, _ = _grid_cfg(np.float32, (16, 16, 16))
    gpu.set_buffer("force_density", np.zeros((3, 16, 16, 16), np.float32))
    stokes_solve_(gpu)
    assert not np.any(gpu.get_buffer("fluid_velocity"))


# ---------------------------------------------------------------- (d) interpolation
@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("geom", [
    dict(L=(8.0, 8.0, 8.0), M=(16, 16, 16), M_G=8, N=500),
    dict(L=(32.0, 32.0, 32.0), M=(64, 64, 64), M_G=8, N=4000),
    dict(L=(12.0, 24.0, 18.0), M=(24, 48, 36), M_G=6, N=1500),
    dict(L=(8.0, 8.0, 8.0), M=(32, 32, 32), M_G=21, N=40),
    dict(L=(24.0, 24.0, 24.0), M=(48, 48, 48), M_G=10, N=2500),  # M_G = 10: two passes per plane
])
def test_interpolation_matches_oracle(T, geom):
    from ForceCouplingMethod import interpolate_velocities_

    rng = np.random.default_rng(8)
    gpu, ora = pair(T, L=geom["L"], R_c=1.0, N=geom["N"], M=geom["M"], M_G=geom["M_G"])
    Y, F = _random_cloud(rng, geom["N"], geom["L"], T, 0.0, 1.0)
    Ys, Fs, orig, cs, ce = _sorted_inputs(ora, Y, F)
    u = rng.normal(size=(3,) + geom["M"]).astype(T)
    gpu.set_buffer("fluid_velocity", u)
    gpu.set_buffer("Y_sorted", Ys.T)
    gpu.set_buffer("original_index", orig)
    interpolate_velocities_(gpu)
    V = gpu.get_buffer("V_output")
    Vref = orc.interpolate(ora, np.ascontiguousarray(np.transpose(u, (0, 3, 2, 1))), Ys, orig).T
    assert rel_l2(V, Vref) <= (2e-6 if T is np.float32 else 1e-13)


def test_interpolation_is_adjoint_of_spreading():
    # test/accuracy/test_interpolate_velocities.jl:8-63: <J u, F> = h^3 <u, S F>
    from ForceCouplingMethod import interpolate_velocities_, spread_forces_

    T = np.float64
    rng = np.random.default_rng(21)
    L, M, N = (16.0, 16.0, 16.0), (32, 32, 32), 200
    gpu, ora = pair(T, L=L, R_c=1.0, N=N, M=M)
    Y, F = _random_cloud(rng, N, L, T, 0.0, 1.0)
    Ys, Fs, orig, cs, ce = _sorted_inputs(ora, Y, F)
    u = rng.normal(size=(3,) + M)
    gpu.set_buffer("Y_sorted", Ys.T)
    gpu.set_buffer("F_sorted", Fs.T)
    gpu.set_buffer("original_index", np.arange(1, N + 1, dtype=np.int32))
    gpu.set_buffer("fluid_velocity", u)
    spread_forces_(gpu)
    interpolate_velocities_(gpu)
    SF = gpu.get_buffer("force_density")
    JU = gpu.get_buffer("V_output")
    h3 = 0.5 ** 3
    assert float((JU * Fs.T).sum()) == pytest.approx(h3 * float((u * SF).sum()), rel=1e-12)


# ---------------------------------------------------------------- (e) correction
@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("case", ["random", "clustered", "sigma_equals_Sigma"])
def test_correction_matches_oracle(T, case):
    # test/cuda/test_gpu_correct.jl:22-32 mirrors the CPU cell list onto the device
    from ForceCouplingMethod import correct_velocities_

    rng = np.random.default_rng(4)
    ratio = 1.0 if case == "sigma_equals_Sigma" else 2.0
    if case == "clustered":
        L, Rc, N = (8.0, 8.0, 8.0), 1.0, 64
        n = np.arange(N)
        Y = np.stack([4 + 0.25 * (n % 4), 4 + 0.25 * ((n // 4) % 4), 4 + 0.25 * (n // 16)]).astype(T)
        F = np.stack([0.1 * (n + 1), -0.2 * (n + 1), 0.3 * (n + 1)]).astype(T)
    else:
        L, Rc, N = (24.0, 24.0, 24.0), 4.0, 3000
        Y, F = _random_cloud(rng, N, L, T, 0.0, 1.0)
    gpu, ora = pair(T, L=L, R_c=Rc, N=N, M=tuple(int(2 * l) for l in L), ratio=ratio)
    Ys, Fs, orig, cs, ce = _sorted_inputs(ora, np.asfortranarray(Y), np.asfortranarray(F))
    V0 = rng.normal(size=(N, 3)).astype(T)
    gpu.set_buffer("Y_sorted", Ys.T)
    gpu.set_buffer("F_sorted", Fs.T)
    gpu.set_buffer("cell_start", cs)
    gpu.set_buffer("cell_end", ce)
    gpu.set_buffer("original_index", orig)
    gpu.set_buffer("V_output", V0.T)
    correct_velocities_(gpu)
    V = gpu.get_buffer("V_output")
    Vref = orc.correct(ora, Ys, Fs, cs, ce, orig, V0.copy()).T
    if case == "sigma_equals_Sigma":
        # Σ = σ: the correction vanishes identically (test_gpu_correct.jl:70-90)
        np.testing.assert_allclose(V, V0.T, rtol=0, atol=100 * np.finfo(T).eps)
        np.testing.assert_allclose(Vref, V0.T, rtol=0, atol=100 * np.finfo(T).eps)
    else:
        assert rel_l2(V - V0.T, Vref - V0.T) <= (1e-5 if T is np.float32 else 1e-12)


def test_float32_correction_against_float64_oracle():
    # Float32 contexts evaluate the pair scalars from cubic pieces fitted to their double-
    # precision values (a 256-interval table over r^2): the correction then agrees with the
    # float64 oracle's on the same (float32-representable) inputs more closely than the float32
    # formulas do, whose terms cancel (the float32 oracle is within ~1e-6 of the float64 one
    # here)
    from ForceCouplingMethod import correct_velocities_

    rng = np.random.default_rng(47)
    L, Rc, N = (24.0, 24.0, 24.0), 4.0, 3000
    Y, F = _random_cloud(rng, N, L, np.float32, 0.0, 1.0)
    M = tuple(int(2 * l) for l in L)
    gpu, ora32 = pair(np.float32, L=L, R_c=Rc, N=N, M=M)
    ora64 = orc.OracleConfig(dtype=np.float64, **kwargs(np.float64, L=L, R_c=Rc, N=N, M=M))
    Ys, Fs, orig, cs, ce = _sorted_inputs(ora32, np.asfortranarray(Y), np.asfortranarray(F))
    gpu.set_buffer("Y_sorted", Ys.T)
    gpu.set_buffer("F_sorted", Fs.T)
    gpu.set_buffer("cell_start", cs)
    gpu.set_buffer("cell_end", ce)
    gpu.set_buffer("original_index", orig)
    gpu.set_buffer("V_output", np.zeros((3, N), np.float32))
    correct_velocities_(gpu)
    V = gpu.get_buffer("V_output").astype(np.float64)
    z = np.zeros((N, 3))
    V64 = orc.correct(ora64, Ys.astype(np.float64), Fs.astype(np.float64), cs, ce, orig, z.copy()).T
    V32 = orc.correct(ora32, Ys, Fs, cs, ce, orig, z.astype(np.float32)).T.astype(np.float64)
    e_gpu, e_ora32 = rel_l2(V, V64), rel_l2(V32, V64)
    print(f"correction vs float64 oracle: GPU float32 {e_gpu:.2e}, oracle float32 {e_ora32:.2e}")
    assert e_gpu <= 3e-7
    assert e_gpu <= e_ora32
    gpu.close()
