"""Device-resident applies replayed from a captured hipGraph (ffcm_set_graph) against the same
applies launched kernel by kernel and against the oracle.

The reference launches every kernel on each mobility! call (ext/ForceCouplingMethodCUDAExt/
mobility.jl:10-60); the captured form must give what those launches give for whatever V, Y and
F hold at replay time: forces and positions changed in place are read by the replay, new
pointers start a new capture, and a scalar change drops the graph.
"""
import numpy as np
import pytest

from helpers import REL_L2_F32, rel_l2
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _oracle_V(T, kw, Y, F):
    cfg = orc.OracleConfig(dtype=T, **kw)
    return orc.mobility(cfg, np.ascontiguousarray(Y.T), np.ascontiguousarray(F.T)).T


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_graph_replays_match_launched_applies(T):
    import torch

    from ForceCouplingMethod import FFCMConfig, mobility_device_
    from ForceCouplingMethod.suspensions import normal_forces, plain_workload

    kw, Y, F = plain_workload(0).typed(T)
    N = Y.shape[1]
    tol = 1e-6 if T == np.float32 else 1e-13  # launched vs replayed: same kernels, same order
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).cuda()  # 3xN column-major = (N, 3) rows
    host = lambda d: d.cpu().numpy().T
    launched = FFCMConfig(gpu_acceleration=True, **kw)
    graphed = FFCMConfig(gpu_acceleration=True, **kw)
    graphed.set_graph(True)
    dY, dF = dev(Y), dev(F)
    dV, dW = torch.empty_like(dY), torch.empty_like(dY)
    mobility_device_(dV, launched, dY, dF)
    ref = host(dV)
    assert rel_l2(ref, _oracle_V(T, kw, Y, F)) <= (REL_L2_F32 if T == np.float32 else 1e-11)
    for k in range(4):  # launched, captured, then replays
        dW.fill_(float("nan"))
        torch.cuda.synchronize()
        mobility_device_(dW, graphed, dY, dF)
        e = rel_l2(host(dW), ref)
        print(f"graph apply {k}: rel L2 vs launched {e:.3e}")
        assert e <= tol
        torch.mean(dW * 2.0).item()  # eager work between replays
    # forces, then positions, changed in place: the replay reads the new values
    for k in range(2):
        Fk = np.asfortranarray(normal_forces(N, 900 + k).astype(T))
        dF.copy_(dev(Fk))
        if k == 1:
            dY[: N // 3, 0] += 0.5
        torch.cuda.synchronize()
        mobility_device_(dW, graphed, dY, dF)
        mobility_device_(dV, launched, dY, dF)
        assert rel_l2(host(dW), host(dV)) <= tol
    Yk = host(dY)
    assert rel_l2(host(dW), _oracle_V(T, kw, np.asfortranarray(Yk), np.asfortranarray(host(dF)))) <= (
        REL_L2_F32 if T == np.float32 else 1e-11)
    # new output buffer: a new capture for the new pointers
    dU = torch.empty_like(dY)
    for k in range(3):
        dU.zero_()
        torch.cuda.synchronize()
        mobility_device_(dU, graphed, dY, dF)
        assert rel_l2(host(dU), host(dV)) <= tol
    launched.close()
    graphed.close()


def test_graph_apply_invalidates_cached_positions():
    import torch

    from ForceCouplingMethod import ArgumentError, FFCMConfig, apply_forces_device_, mobility_device_, set_positions_device_
    from ForceCouplingMethod.suspensions import plain_workload

    T = np.float32
    kw, Y, F = plain_workload(0).typed(T)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
    cfg = FFCMConfig(gpu_acceleration=True, **kw)
    cfg.set_graph(True)
    dY, dF = dev(Y), dev(F)
    dV = torch.empty_like(dY)
    for _ in range(3):
        mobility_device_(dV, cfg, dY, dF)
    set_positions_device_(cfg, dY)
    apply_forces_device_(dV, cfg, dF)
    mobility_device_(dV, cfg, dY, dF)  # a replay rebuilds the cell list, as a launched apply does
    with pytest.raises(ArgumentError, match="set_positions"):
        apply_forces_device_(dV, cfg, dF)
    cfg.set_graph(False)
    cfg.close()


def test_graph_dropped_on_scalar_change():
    import ctypes

    import torch

    from ForceCouplingMethod import FFCMConfig, _lib, mobility_device_
    from ForceCouplingMethod.config import derive_scalars, params_of
    from ForceCouplingMethod.suspensions import plain_workload

    T = np.float32
    kw, Y, F = plain_workload(0).typed(T)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
    cfg = FFCMConfig(gpu_acceleration=True, **kw)
    cfg.set_graph(True)
    dY, dF = dev(Y), dev(F)
    dV = torch.empty_like(dY)
    for _ in range(3):
        mobility_device_(dV, cfg, dY, dF)
    V1 = dV.cpu().numpy().T.copy()
    # twice the viscosity (ffcm_set_scalars): the captured kernels hold the old scalars, so the
    # graph is dropped, and the velocity halves
    p = params_of(derive_scalars(**dict(kw, viscosity=T(2))))
    assert _lib.lib().ffcm_set_scalars(cfg.ctx, ctypes.byref(p)) == 0, _lib.last_error(cfg.ctx)
    for _ in range(3):
        mobility_device_(dV, cfg, dY, dF)
        assert rel_l2(dV.cpu().numpy().T, V1 / 2) <= 1e-6
    cfg.close()
