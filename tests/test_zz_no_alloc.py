"""The allocation-free entry point ffcm_mobility_device_noalloc (no reference counterpart: a
variant of ffcm_mobility_device for callers that cannot allocate device memory) against the
context-based ffcm_mobility_device, through the C ABI, on two calls in a row with the second
one smaller; and its size limit."""
import ctypes

import numpy as np
import pytest

from helpers import REL_L2_F32, rel_l2
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _lib():
    from ForceCouplingMethod import _lib as L

    lib = L.lib()
    vp = ctypes.c_void_p
    lib.ffcm_mobility_device_noalloc.argtypes = [ctypes.POINTER(L.Params), vp, vp, vp, ctypes.c_int64, vp]
    lib.ffcm_mobility_device_noalloc.restype = ctypes.c_int
    lib.ffcm_noalloc_last_error.argtypes = []
    lib.ffcm_noalloc_last_error.restype = ctypes.c_char_p
    lib.ffcm_noalloc_scratch_bytes.argtypes = []
    lib.ffcm_noalloc_scratch_bytes.restype = ctypes.c_size_t
    return L, lib


def _case(N, L, M, seed):
    from ForceCouplingMethod.suspensions import normal_forces, rsa_positions

    T = np.float32
    Y, _ = rsa_positions(N, L, seed)
    F = normal_forces(N, seed + 1)
    kw = dict(L=(T(L),) * 3, R_c=T(4), N=N, kernel_widths_ratio=T(2), num_grid_points=(M,) * 3, M_G=8,
              viscosity=T(1))
    return kw, np.asfortranarray(Y.astype(T)), np.asfortranarray(F.astype(T))


def test_noalloc_matches_context_entry_point_twice_smaller_second():
    import torch

    from ForceCouplingMethod import FFCMConfig, mobility_device_
    from ForceCouplingMethod.config import derive_scalars, params_of

    L, lib = _lib()
    stream = torch.cuda.current_stream().cuda_stream
    for N, Lbox, M, seed in ((3000, 48.0, 96, 71), (512, 32.0, 64, 72)):
        kw, Y, F = _case(N, Lbox, M, seed)
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
        dY, dF = dev(Y), dev(F)
        dV_ref = torch.empty_like(dY)
        cfg = FFCMConfig(gpu_acceleration=True, **kw)
        mobility_device_(dV_ref, cfg, dY, dF)
        cfg.close()
        d = derive_scalars(**kw)
        p = params_of(d)
        dV = torch.full_like(dY, float("nan"))
        torch.cuda.synchronize()
        rc = lib.ffcm_mobility_device_noalloc(ctypes.byref(p), dV.data_ptr(), dY.data_ptr(), dF.data_ptr(), N, stream)
        assert rc == 0, lib.ffcm_noalloc_last_error().decode()
        torch.cuda.synchronize()
        V, Vref = dV.cpu().numpy().T, dV_ref.cpu().numpy().T
        e = rel_l2(V, Vref)
        print(f"noalloc N={N} {M}^3: rel L2 vs ffcm_mobility_device {e:.3e}")
        assert np.isfinite(V).all()
        assert e <= 1e-6
        Vo = orc.mobility(orc.OracleConfig(dtype=np.float32, **kw), np.ascontiguousarray(Y.T),
                          np.ascontiguousarray(F.T)).T
        assert rel_l2(V, Vo) <= REL_L2_F32


def test_noalloc_rejects_configurations_beyond_its_scratch():
    import torch

    from ForceCouplingMethod.config import derive_scalars, params_of

    L, lib = _lib()
    assert lib.ffcm_noalloc_scratch_bytes() == 128 << 20
    kw, Y, F = _case(1000, 140.0, 256, 73)  # a 256^3 grid needs about 0.6 GB
    dsc = derive_scalars(**kw)
    p = params_of(dsc)
    d = torch.zeros(3 * 1000, dtype=torch.float32, device="cuda")
    rc = lib.ffcm_mobility_device_noalloc(ctypes.byref(p), d.data_ptr(), d.data_ptr(), d.data_ptr(), 1000, None)
    assert rc == L.FFCM_ERR_ARGUMENT
    assert "bytes of device scratch" in lib.ffcm_noalloc_last_error().decode()
