"""Pipeline-level GPU checks: the production cell ranges of a full apply, the whole neighbour
map, run-to-run determinism of each deterministic stage, and a clustered suspension at phi = 0.2
against the oracle.

The full apply builds cell_start / cell_end in the fused gather-and-bin kernel (not the
per-step cell_offsets kernel); they must equal the reference's stable counting sort
(src/cell_list.jl:64-93) bit for bit. Determinism: the cell list, the FFT Stokes solve, the
interpolation and the correction are atomic-free and must repeat bit for bit; the spread ranks
each particle in its tile with an atomic, so its summation order (and last bits) may vary
between runs (checked to rel L2 1e-6).
"""
import numpy as np
import pytest

from helpers import REL_L2_F32, neighbor_map, oracle_rows, pair, rel_l2
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def test_full_apply_cell_ranges_are_bit_exact():
    from ForceCouplingMethod import FFCMConfig, mobility_
    from ForceCouplingMethod.suspensions import plain_workload

    T = np.float32
    for idx in (0, 1):
        kw, Y, F = plain_workload(idx).typed(T)
        cfg = FFCMConfig(gpu_acceleration=True, **kw)
        mobility_(np.empty(Y.shape, T, order="F"), cfg, Y, F)
        ora = orc.OracleConfig(dtype=T, **kw)
        Yw = orc.wrap_positions(ora, oracle_rows(Y))
        h = orc.assign_cells(ora, Yw)
        orig, cs, ce = orc.build_cell_list(h, ora.total_cells)
        assert np.array_equal(cfg.get_buffer("cell_hash"), h)
        assert np.array_equal(cfg.get_buffer("original_index"), orig)
        assert np.array_equal(cfg.get_buffer("cell_start"), cs)
        assert np.array_equal(cfg.get_buffer("cell_end"), ce)
        Ys, Fs = orc.gather(ora, orig, Yw, oracle_rows(F))
        assert np.array_equal(cfg.get_buffer("Y_sorted"), Ys.T)
        assert np.array_equal(cfg.get_buffer("F_sorted"), Fs.T)
        cfg.close()


@pytest.mark.parametrize("L,Rc", [((8.0, 12.0, 16.0), 2.0), ((30.0, 30.0, 30.0), 4.0)])
def test_whole_neighbor_map(L, Rc):
    gpu, ora = pair(np.float32, L=L, R_c=Rc, N=8, M=tuple(int(2 * x) for x in L))
    m = gpu.get_buffer("neighbor_map")
    assert m.shape == (13 * ora.total_cells,)
    assert np.array_equal(m, neighbor_map(ora.num_cells))
    gpu.close()


def test_deterministic_stages_repeat_bit_for_bit():
    from ForceCouplingMethod import (FFCMConfig, correct_velocities_, interpolate_velocities_, mobility_,
                                     spread_forces_, stokes_solve_)
    from ForceCouplingMethod.suspensions import plain_workload

    T = np.float32
    kw, Y, F = plain_workload(1).typed(T)
    cfg = FFCMConfig(gpu_acceleration=True, **kw)
    mobility_(np.empty(Y.shape, T, order="F"), cfg, Y, F)
    cells = [cfg.get_buffer(b) for b in ("cell_hash", "original_index", "cell_start", "cell_end", "Y_sorted")]
    mobility_(np.empty(Y.shape, T, order="F"), cfg, Y, F)
    for b, ref in zip(("cell_hash", "original_index", "cell_start", "cell_end", "Y_sorted"), cells):
        assert np.array_equal(cfg.get_buffer(b), ref), b
    # spread twice: equal to rounding order
    spread_forces_(cfg)
    f1 = cfg.get_buffer("force_density")
    spread_forces_(cfg)
    assert rel_l2(cfg.get_buffer("force_density"), f1) <= 1e-6
    # FFT solve, interpolation and correction on fixed inputs: bit for bit
    cfg.set_buffer("force_density", f1)
    stokes_solve_(cfg)
    u1 = cfg.get_buffer("fluid_velocity")
    cfg.set_buffer("force_density", f1)
    stokes_solve_(cfg)
    assert np.array_equal(cfg.get_buffer("fluid_velocity"), u1)
    interpolate_velocities_(cfg)
    v1 = cfg.get_buffer("V_output")
    interpolate_velocities_(cfg)
    assert np.array_equal(cfg.get_buffer("V_output"), v1)
    cfg.set_buffer("V_output", np.zeros_like(v1))
    correct_velocities_(cfg)
    c1 = cfg.get_buffer("V_output")
    cfg.set_buffer("V_output", np.zeros_like(v1))
    correct_velocities_(cfg)
    assert np.array_equal(cfg.get_buffer("V_output"), c1)
    cfg.close()


def test_clustered_phi_020_suspension_matches_oracle():
    # 2^18 spheres at phi = 0.20 in L = 176 (320^3 grid, h = 0.55), half of them in 16 Gaussian
    # clusters of std 8 (10,000 draws before a uniform fallback): the dense regime of the
    # 2^20-particle benchmark suspension at a size the oracle finishes in seconds on 16 threads
    from ForceCouplingMethod import FFCMConfig, mobility_
    from ForceCouplingMethod.suspensions import normal_forces, rsa_positions

    T = np.float32
    N, L, M = 262144, 176.0, 320
    Y, fb = rsa_positions(N, L, 0xC1A5, n_clusters=16, n_clustered=N // 2, cluster_std=8.0,
                          max_cluster_tries=10000)
    print(f"clustered suspension: {fb} of {N // 2} cluster draws fell back to uniform")
    assert fb < N // 4
    F = normal_forces(N, 0xC1A6)
    kw = dict(L=(T(L),) * 3, R_c=T(4), N=N, kernel_widths_ratio=T(2), num_grid_points=(M,) * 3, M_G=8,
              viscosity=T(1))
    Y32, F32 = np.asfortranarray(Y.astype(T)), np.asfortranarray(F.astype(T))
    cfg = FFCMConfig(gpu_acceleration=True, **kw)
    V = mobility_(np.empty((3, N), T, order="F"), cfg, Y32, F32)
    cfg.close()
    orc.set_threads(16, 16)
    try:
        Vo = orc.mobility(orc.OracleConfig(dtype=T, **kw), np.ascontiguousarray(Y32.T),
                          np.ascontiguousarray(F32.T)).T
    finally:
        orc.set_threads(1, 1)
    e = rel_l2(V, Vo)
    print(f"clustered phi=0.20, N={N}, {M}^3: rel L2 vs CPU f32 {e:.3e}")
    assert e <= REL_L2_F32
