"""The iterative-solver operator with cached positions and the device-resident entry points
against the oracle.

FFCMMobility (src/mobility.jl:88-162) keeps its positions' cell list in the config between
applies (ffcm_set_positions / ffcm_apply_forces); the reference re-sorts on every mul!
(docs/src/devdocs/gpu-architecture.md:159-161), so each cached apply must equal
mobility!(V, config, Y, F) for the operator's Y. The device-resident forms take torch tensors
already in HBM (the device branch of _stage_mobility_io!, ext/ForceCouplingMethodCUDAExt/
mobility.jl:1-8).
"""
import numpy as np
import pytest

from helpers import REL_L2_F32, rel_l2
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _oracle_V(T, kw, Y, F):
    cfg = orc.OracleConfig(dtype=T, **kw)
    return orc.mobility(cfg, np.ascontiguousarray(Y.T), np.ascontiguousarray(F.T)).T


@pytest.mark.parametrize("T,tol", [(np.float32, REL_L2_F32), (np.float64, 1e-11)])
def test_cached_operator_matches_oracle_over_many_applies(T, tol):
    from ForceCouplingMethod import FFCMConfig, FFCMMobility, mobility_, mul_
    from ForceCouplingMethod.suspensions import normal_forces, plain_workload

    kw, Y, F = plain_workload(0).typed(T)
    N = Y.shape[1]
    if T is np.float64:
        with pytest.warns(UserWarning):
            cfg = FFCMConfig(gpu_acceleration=True, **kw)
    else:
        cfg = FFCMConfig(gpu_acceleration=True, **kw)
    M = FFCMMobility(cfg, Y)
    v = np.empty(3 * N, T)
    for k in range(3):  # the second and third applies reuse the cell list of the first
        Fk = np.asfortranarray(normal_forces(N, 300 + k).astype(T))
        mul_(v, M, Fk.reshape(-1, order="F"))
        e = rel_l2(v.reshape(3, N, order="F"), _oracle_V(T, kw, Y, Fk))
        print(f"cached apply {k}: rel L2 vs CPU {e:.3e}")
        assert e <= tol
    # another caller of the same config replaces the cell list; the operator notices
    Y2 = np.asfortranarray(Y + T(3.25))
    V2 = mobility_(np.empty((3, N), T, order="F"), cfg, Y2, F)
    assert rel_l2(V2, _oracle_V(T, kw, Y2, F)) <= tol
    mul_(v, M, F.reshape(-1, order="F"))
    assert rel_l2(v.reshape(3, N, order="F"), _oracle_V(T, kw, Y, F)) <= tol
    # 5-argument mul!: v = alpha M f + beta v
    w0 = np.random.default_rng(5).normal(size=3 * N).astype(T)
    w = w0.copy()
    mul_(w, M, F.reshape(-1, order="F"), 2.0, -0.5)
    ref = 2.0 * _oracle_V(T, kw, Y, F).reshape(-1, order="F") - 0.5 * w0
    assert rel_l2(w, ref) <= 4 * tol
    cfg.close()


def test_apply_forces_needs_cached_positions():
    from ForceCouplingMethod import ArgumentError, FFCMConfig, apply_forces_, mobility_, set_positions_
    from ForceCouplingMethod.suspensions import plain_workload

    kw, Y, F = plain_workload(0).typed(np.float32)
    cfg = FFCMConfig(gpu_acceleration=True, **kw)
    out = np.empty(Y.shape, np.float32, order="F")
    with pytest.raises(ArgumentError, match="set_positions"):
        apply_forces_(out, cfg, F)
    set_positions_(cfg, Y)
    apply_forces_(out, cfg, F)
    mobility_(np.empty_like(out), cfg, Y, F)  # replaces the cached cell list
    with pytest.raises(ArgumentError, match="set_positions"):
        apply_forces_(out, cfg, F)
    cfg.close()


def test_device_entry_points_match_oracle():
    import torch

    from ForceCouplingMethod import (FFCMConfig, apply_forces_device_, mobility_device_, set_positions_device_)
    from ForceCouplingMethod.suspensions import normal_forces, plain_workload

    T = np.float32
    kw, Y, F = plain_workload(0).typed(T)
    N = Y.shape[1]
    cfg = FFCMConfig(gpu_acceleration=True, **kw)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).cuda()  # 3xN column-major = (N, 3) rows
    dY, dF = dev(Y), dev(F)
    dV = torch.empty_like(dY)
    mobility_device_(dV, cfg, dY, dF)
    e = rel_l2(dV.cpu().numpy().T, _oracle_V(T, kw, Y, F))
    print(f"device apply: rel L2 vs CPU f32 {e:.3e}")
    assert e <= REL_L2_F32
    set_positions_device_(cfg, dY)
    for k in range(2):
        Fk = np.asfortranarray(normal_forces(N, 700 + k).astype(T))
        apply_forces_device_(dV, cfg, dev(Fk))
        assert rel_l2(dV.cpu().numpy().T, _oracle_V(T, kw, Y, Fk)) <= REL_L2_F32
    cfg.close()


def test_cached_operator_at_65536_particles():
    # the position-cached apply on the 65,536-particle / 256^3 suspension against the uncached
    # apply and the oracle
    from ForceCouplingMethod import FFCMConfig, apply_forces_, mobility_, set_positions_
    from ForceCouplingMethod.suspensions import normal_forces, plain_workload

    T = np.float32
    kw, Y, F = plain_workload(1).typed(T)
    N = Y.shape[1]
    cfg = FFCMConfig(gpu_acceleration=True, **kw)
    G = np.asfortranarray(normal_forces(N, 77).astype(T))
    V = mobility_(np.empty((3, N), T, order="F"), cfg, Y, G)
    set_positions_(cfg, Y)
    Vc = apply_forces_(np.empty((3, N), T, order="F"), cfg, G)
    assert rel_l2(Vc, V) <= 1e-6
    orc.set_threads(16, 16)
    try:
        e = rel_l2(Vc, _oracle_V(T, kw, Y, G))
    finally:
        orc.set_threads(1, 1)
    assert e <= REL_L2_F32
    cfg.close()


def test_cached_operator_sees_positions_changed_in_place():
    from ForceCouplingMethod import FFCMConfig, FFCMMobility
    from ForceCouplingMethod.suspensions import plain_workload

    T = np.float32
    kw, Y, F = plain_workload(0).typed(T)
    cfg = FFCMConfig(gpu_acceleration=True, **kw)
    Yop = np.asfortranarray(Y.copy())
    M = FFCMMobility(cfg, Yop)
    V1 = M @ F
    M.Y[0, : Y.shape[1] // 2] += T(0.75)  # half the particles move in place (the reference reads Y per mul!)
    V2 = M @ F
    assert rel_l2(V2, _oracle_V(T, kw, np.asfortranarray(M.Y), F)) <= REL_L2_F32
    assert rel_l2(V2, V1) > 1e-3
    cfg.close()
