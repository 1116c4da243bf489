"""Column-streamed interpolation on the GPU (the kernels chosen where particle stencils cover the
grid densely): the ring filled by LDS DMA (FFCM_INTERP=dma, the default there) and the ring
filled through registers (FFCM_INTERP=column), each against the per-particle kernel
(FFCM_INTERP=wave) and against the oracle's interpolate (src/interpolate.jl:23-136).

Cases cover the block layouts of the column kernels: columns visited in 8 x 4 patches and in
rows, blocks holding one tile layer, whole columns with an even and an odd number of layers,
stencils wrapping around every face, and a tile holding more particles (600) than either kernel
stages in LDS (512 and 64). The kernels sum in different orders, so they agree to rel L2 1e-6;
the oracle bound is the one of the interpolation step test.
"""
import numpy as np
import pytest

from helpers import oracle_rows, pair, rel_l2
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SAME = 1e-6


def _sorted_inputs(ora, Y):
    Yw = orc.wrap_positions(ora, oracle_rows(Y))
    h = orc.assign_cells(ora, Yw)
    orig, _, _ = orc.build_cell_list(h, ora.total_cells)
    Ys, _ = orc.gather(ora, orig, Yw, np.zeros_like(Yw))
    return Ys, orig


def _interpolate(gpu, u, Ys, orig, monkeypatch, kernel, segs=None, expect=None):
    from ForceCouplingMethod import interpolate_velocities_

    monkeypatch.setenv("FFCM_INTERP", kernel)
    if segs is None:
        monkeypatch.delenv("FFCM_INTERP_SEGS", raising=False)
    else:
        monkeypatch.setenv("FFCM_INTERP_SEGS", str(segs))
    assert gpu.info()["interp"] == (expect or kernel)
    gpu.set_buffer("fluid_velocity", u)
    gpu.set_buffer("Y_sorted", Ys.T)
    gpu.set_buffer("original_index", orig)
    interpolate_velocities_(gpu)
    return gpu.get_buffer("V_output").copy()


CASES = [
    # (name, L, M, N, FFCM_INTERP_SEGS)
    ("patches_one_layer_blocks", (64.0, 32.0, 24.0), (128, 64, 48), 3000, None),
    ("rows_one_layer_blocks", (24.0, 40.0, 32.0), (48, 80, 64), 4000, None),
    ("whole_columns_even_layers", (16.0, 16.0, 32.0), (32, 32, 64), 2500, 1),
    ("whole_columns_odd_layers", (16.0, 16.0, 20.0), (32, 32, 40), 2000, 1),
    ("three_segments", (24.0, 16.0, 28.0), (48, 32, 56), 3000, 3),
]


COLUMN_KERNELS = ["dma", "column"]


@pytest.mark.parametrize("kernel", COLUMN_KERNELS)
@pytest.mark.parametrize("name,L,M,N,segs", CASES, ids=[c[0] for c in CASES])
def test_column_interpolation_matches_wave_and_oracle(name, L, M, N, segs, kernel, monkeypatch):
    T = np.float32
    rng = np.random.default_rng(abs(hash(name)) % 2**32)
    gpu, ora = pair(T, L=L, R_c=1.0, N=N, M=M)
    Y = np.empty((3, N), dtype=T, order="F")
    for i in range(3):  # positions in [-L, 2L): every stencil wraps somewhere
        Y[i] = rng.uniform(-L[i], 2 * L[i], size=N).astype(T)
    u = rng.normal(size=(3,) + M).astype(T)
    Ys, orig = _sorted_inputs(ora, Y)
    Vc = _interpolate(gpu, u, Ys, orig, monkeypatch, kernel, segs)
    Vw = _interpolate(gpu, u, Ys, orig, monkeypatch, "wave")
    assert rel_l2(Vc, Vw) <= SAME
    Vref = orc.interpolate(ora, np.ascontiguousarray(np.transpose(u, (0, 3, 2, 1))), Ys, orig).T
    assert rel_l2(Vc, Vref) <= 2e-6
    gpu.close()


@pytest.mark.parametrize("kernel", COLUMN_KERNELS)
def test_column_interpolation_tile_beyond_lds_staging(kernel, monkeypatch):
    # 600 particles whose stencils start in tile (0, 0, 0) (16 x 16 x 8 points, h = 0.5), plus a
    # uniform background: the first layer of column (0, 0) holds more than either kernel stages
    T = np.float32
    L, M = (16.0, 16.0, 8.0), (32, 32, 16)
    rng = np.random.default_rng(600)
    dense = np.stack([rng.uniform(2.0, 5.9, 600), rng.uniform(2.0, 5.9, 600), rng.uniform(2.6, 3.4, 600)])
    back = np.stack([rng.uniform(0, l, 400) for l in L])
    Y = np.asfortranarray(np.concatenate([dense, back], axis=1).astype(T))
    N = Y.shape[1]
    gpu, ora = pair(T, L=L, R_c=1.0, N=N, M=M)
    # stencil low corners (round(y / h) - 4) of the dense particles lie in [0, 16) x [0, 16) x [0, 8)
    low = np.rint(dense / 0.5) - 4
    assert low.min() >= 0 and low[:2].max() < 16 and low[2].max() < 8
    u = rng.normal(size=(3,) + M).astype(T)
    Ys, orig = _sorted_inputs(ora, Y)
    Vc = _interpolate(gpu, u, Ys, orig, monkeypatch, kernel, 1)
    Vw = _interpolate(gpu, u, Ys, orig, monkeypatch, "wave")
    assert rel_l2(Vc, Vw) <= SAME
    Vref = orc.interpolate(ora, np.ascontiguousarray(np.transpose(u, (0, 3, 2, 1))), Ys, orig).T
    assert rel_l2(Vc, Vref) <= 2e-6
    gpu.close()


def test_dense_mobility_uses_column_kernel_and_matches_oracle():
    # 4 stencil coverings of the grid (the column kernel's default range): the full apply
    # against the oracle's Float32 mobility
    from ForceCouplingMethod import FFCMConfig, mobility_
    from ForceCouplingMethod.suspensions import normal_forces

    T = np.float32
    L, M = (20.0, 20.0, 20.0), (32, 32, 32)
    N = 256
    rng = np.random.default_rng(44)
    Y = np.asfortranarray(rng.uniform(0, 20.0, size=(3, N)).astype(T))
    F = np.asfortranarray(normal_forces(N, 45).astype(T))
    assert N * 512 / (32 ** 3) >= 3.0
    kw = dict(L=tuple(T(x) for x in L), R_c=T(4), N=N, kernel_widths_ratio=T(2), num_grid_points=M, M_G=8,
              viscosity=T(1))
    cfg = FFCMConfig(gpu_acceleration=True, **kw)
    assert cfg.info()["interp"] == "dma"
    V = mobility_(np.empty_like(F), cfg, Y, F)
    cfg.close()
    Vo = orc.mobility(orc.OracleConfig(dtype=T, **kw), np.ascontiguousarray(Y.T), np.ascontiguousarray(F.T)).T
    assert rel_l2(V, Vo) <= 1e-5


@pytest.mark.parametrize("kernel", COLUMN_KERNELS)
def test_thin_grid_falls_back_from_column_kernel(kernel, monkeypatch):
    # Mz = 8: the column kernels' rings would reach planes past the grid's second period, so
    # the plan refuses them and the per-particle kernel serves even when FFCM_INTERP asks for a
    # column one (128 x 128 x 8 grid, 16 x 16 x 8 tiles)
    T = np.float32
    L, M, N = (64.0, 64.0, 4.0), (128, 128, 8), 3000
    rng = np.random.default_rng(808)
    gpu, ora = pair(T, L=L, R_c=1.0, N=N, M=M)
    assert gpu.info()["tile"] == (16, 16, 8)
    Y = np.asfortranarray(np.stack([rng.uniform(-L[i], 2 * L[i], N) for i in range(3)]).astype(T))
    u = rng.normal(size=(3,) + M).astype(T)
    Ys, orig = _sorted_inputs(ora, Y)
    Vc = _interpolate(gpu, u, Ys, orig, monkeypatch, kernel, expect="wave")
    Vref = orc.interpolate(ora, np.ascontiguousarray(np.transpose(u, (0, 3, 2, 1))), Ys, orig).T
    assert rel_l2(Vc, Vref) <= 2e-6
    gpu.close()
