"""Several right-hand sides per system and the batched ensemble on the GPU.

Semantics (no reference function): every (system, right-hand side) result equals a separate
mobility! on that system's FFCMConfig (src/mobility.jl:23-45). Checked against the plain GPU
apply, which is itself parity-tested against the oracle, and directly against the oracle on
small systems. Sums in spreading are order-dependent, so equality is to rel L2 1e-6.
"""
import numpy as np
import pytest

from helpers import REL_L2_F32, rel_l2
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SAME = 1e-6  # GPU vs GPU, different launch grouping


def _plain(kw, Y, F):
    from ForceCouplingMethod import FFCMConfig, mobility_

    cfg = FFCMConfig(gpu_acceleration=True, **kw)
    V = mobility_(np.empty_like(F), cfg, Y, F)
    cfg.close()
    return V


def test_multi_rhs_equals_separate_applies():
    from ForceCouplingMethod import FFCMConfig, mobility_, mobility_multi_
    from ForceCouplingMethod.suspensions import normal_forces, plain_workload

    w = plain_workload(0)
    kw, Y, _ = w.typed(np.float32)
    N = Y.shape[1]
    R = 5
    F = np.asfortranarray(np.transpose(normal_forces(N, 0xB1, n_rhs=R), (1, 2, 0)).astype(np.float32))
    cfg = FFCMConfig(gpu_acceleration=True, max_rhs=6, **kw)
    V = mobility_multi_(np.empty((3, N, R), np.float32, order="F"), cfg, Y, F)
    for r in range(R):
        Vr = mobility_(np.empty((3, N), np.float32, order="F"), cfg, Y, np.asfortranarray(F[:, :, r]))
        assert rel_l2(V[:, :, r], Vr) <= SAME, r
    # fewer right-hand sides than the capacity, and the oracle for one of them
    V2 = mobility_multi_(np.empty((3, N, 2), np.float32, order="F"), cfg, Y, np.asfortranarray(F[:, :, :2]))
    assert rel_l2(V2, V[:, :, :2]) <= SAME
    ocfg = orc.OracleConfig(dtype=np.float32, **kw)
    Vo = orc.mobility(ocfg, np.ascontiguousarray(Y.T), np.ascontiguousarray(F[:, :, 3].T)).T
    assert rel_l2(V[:, :, 3], Vo) <= REL_L2_F32
    with pytest.raises(Exception):
        mobility_multi_(np.empty((3, N, 7), np.float32, order="F"), cfg, Y,
                        np.zeros((3, N, 7), np.float32, order="F"))


def _small_ensemble(n_systems, R, seed):
    from ForceCouplingMethod.suspensions import normal_forces, rsa_positions, smooth5_at_least

    rng = np.random.default_rng(seed)
    systems = []
    for k in range(n_systems):
        N = int(np.exp(rng.uniform(np.log(150), np.log(3000))))
        L = (N * 4.0 * np.pi / 3.0 / 0.1) ** (1 / 3)
        M = smooth5_at_least(int(np.ceil(L / 0.55)))
        Y, _ = rsa_positions(N, L, 0x5000 + k)
        F = np.transpose(normal_forces(N, 0x6000 + k, n_rhs=R), (1, 2, 0)) if R > 1 else \
            normal_forces(N, 0x6000 + k)[:, :, None]
        kw = dict(L=(np.float32(L),) * 3, R_c=np.float32(4), N=N, kernel_widths_ratio=np.float32(2),
                  num_grid_points=(M, M, M), M_G=8, viscosity=np.float32(1))
        systems.append((kw, np.asfortranarray(Y, np.float32), np.asfortranarray(F, np.float32)))
    return systems


@pytest.mark.parametrize("mode", ["grouped", "lanes"])
def test_batch_matches_per_system_applies_and_oracle(mode, monkeypatch):
    # both executions: one grouped launch per pipeline step over all systems, and stream lanes
    from ForceCouplingMethod import FFCMBatch

    monkeypatch.setenv("FFCM_BATCH_MODE", mode)
    R = 4
    systems = _small_ensemble(7, R, 11)
    batch = FFCMBatch([s[0] for s in systems], n_rhs=R)
    V = [np.empty((3, s[0]["N"], R), np.float32, order="F") for s in systems]
    batch.mobility_(V, [s[1] for s in systems], [s[2] for s in systems])
    info = batch.info()
    assert info["execution"] == mode and info["graph_nodes"] > 0, info
    assert 1 <= info["lanes"] <= info["lanes_allocated"], info
    if mode == "grouped":  # per lane, one node per step and kernel instance, whatever the number of systems
        assert info["graph_nodes"] < 80 * info["lanes"], info
    for k, (kw, Y, F) in enumerate(systems):
        for r in range(R):
            Vr = _plain(kw, Y, np.asfortranarray(F[:, :, r]))
            assert rel_l2(V[k][:, :, r], Vr) <= SAME, (k, r)
    kw, Y, F = systems[2]
    Vo = orc.mobility(orc.OracleConfig(dtype=np.float32, **kw), np.ascontiguousarray(Y.T),
                      np.ascontiguousarray(F[:, :, 1].T)).T
    assert rel_l2(V[2][:, :, 1], Vo) <= REL_L2_F32
    # a second apply replays the captured graph; new forces give new results
    F2 = [np.asfortranarray(-2 * s[2]) for s in systems]
    V2 = [np.empty_like(v) for v in V]
    batch.mobility_(V2, [s[1] for s in systems], F2)
    for k in range(len(systems)):
        assert rel_l2(V2[k], -2 * V[k]) <= SAME


@pytest.mark.parametrize("mode", ["grouped", "lanes"])
def test_batch_device_pointers_and_recapture(mode, monkeypatch):
    import torch

    from ForceCouplingMethod import FFCMBatch

    monkeypatch.setenv("FFCM_BATCH_MODE", mode)
    R = 3
    systems = _small_ensemble(5, R, 12)
    batch = FFCMBatch([s[0] for s in systems], n_rhs=R)
    Vh = [np.empty((3, s[0]["N"], R), np.float32, order="F") for s in systems]
    batch.mobility_(Vh, [s[1] for s in systems], [s[2] for s in systems])

    def dev(a):  # Fortran 3 x N (x R) -> flat device buffer of the same memory order
        return torch.from_numpy(np.ravel(a, order="F").copy()).to("cuda:0")

    for attempt in range(2):  # two sets of device buffers: the second call re-captures
        Yd = [dev(s[1]) for s in systems]
        Fd = [dev(s[2]) for s in systems]
        Vd = [torch.empty_like(f) for f in Fd]
        torch.cuda.synchronize()
        batch.mobility_device_(Vd, Yd, Fd)
        for k, s in enumerate(systems):
            got = Vd[k].cpu().numpy().reshape(Vh[k].shape, order="F")
            assert rel_l2(got, Vh[k]) <= SAME, (attempt, k)


def test_full_ensemble_64_systems_16_rhs_spot_checks():
    # 64 systems, N log-uniform in [1e3, 5e4], 5-smooth grids 64..240, 16 right-hand sides:
    # every system's right-hand sides 0 and 15 against separate plain applies.
    from ForceCouplingMethod import FFCMBatch
    from ForceCouplingMethod.suspensions import ensemble_workload

    ens = ensemble_workload()
    typed = [s.typed(np.float32) for s in ens]
    batch = FFCMBatch([t[0] for t in typed], n_rhs=16)
    V = [np.empty_like(t[2]) for t in typed]
    batch.mobility_(V, [t[1] for t in typed], [t[2] for t in typed])
    worst = 0.0
    for k, (kw, Y, F) in enumerate(typed):
        for r in (0, 15):
            worst = max(worst, rel_l2(V[k][:, :, r], _plain(kw, Y, np.asfortranarray(F[:, :, r]))))
    info = batch.info()
    print(f"ensemble: worst rel L2 vs separate applies {worst:.3e}; {info}")
    assert worst <= SAME
    if info["execution"] == "grouped":
        assert info["graph_nodes"] < 300, info
