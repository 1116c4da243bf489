"""The reference's seams driven call for call through the C ABI, as the unmodified reference
makes them with this backend's package extension loaded: _assemble_gpu_buffers with its nine
arguments (src/config.jl:495-502), then mobility!'s generic body (src/mobility.jl:35-43) with
every L1 barrier receiving the scalars the reference passes it, and the fused mobility! method
(ffcm_set_scalars once, then ffcm_mobility). Both must match the oracle and the ffcm_create
context."""
import ctypes

import numpy as np
import pytest

from helpers import REL_L2_F32, neighbor_map, rel_l2
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

IP = ctypes.POINTER(ctypes.c_int32)
DP = ctypes.POINTER(ctypes.c_double)


def _i3(v):
    return (ctypes.c_int32 * 3)(*[int(x) for x in v])


def _d3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


def _assemble(L, d, T):
    from ForceCouplingMethod import _lib

    nm = neighbor_map(d["num_cells"])
    C = int(np.prod(d["num_cells"]))
    ctx = ctypes.c_void_p()
    rc = L.ffcm_assemble(ctypes.byref(ctx), 0 if T is np.float32 else 1, d["N"], _i3(d["num_grid_points"]),
                         int(d["M_G"]), C, d["k_x"].ctypes.data, d["k_y"].ctypes.data, d["k_z"].ctypes.data,
                         nm.ctypes.data_as(IP))
    assert rc == 0, _lib.last_error(None)
    return ctx


def _check(L, ctx, rc):
    from ForceCouplingMethod import _lib

    assert rc == 0, _lib.last_error(ctx)


def _put(L, ctx, buf, a):
    _check(L, ctx, L.ffcm_set_buffer(ctx, buf, a.ctypes.data, a.nbytes))


def _get(L, ctx, buf, shape, T):
    out = np.empty(shape, dtype=T, order="F")
    _check(L, ctx, L.ffcm_get_buffer(ctx, buf, out.ctypes.data, out.nbytes))
    return out


def _generic_mobility(L, ctx, d, Y, F):
    """mobility!(V, config, Y, F) of src/mobility.jl:23-45 with the HIP barrier methods."""
    from ForceCouplingMethod import _lib

    B = _lib.BUFFER_IDS
    N, T = d["N"], d["T"]
    M, MG = _i3(d["num_grid_points"]), int(d["M_G"])
    sig, Sig, h, ih = (float(d[k]) for k in ("sigma", "Sigma", "h", "inv_h"))
    _put(L, ctx, B["Y_input"], Y)  # _stage_mobility_io!
    _put(L, ctx, B["F_input"], F)
    _check(L, ctx, L.ffcm_wrap_positions_kernel(ctx, B["Y_wrapped"], B["Y_input"], _d3(d["L"])))
    _check(L, ctx, L.ffcm_assign_cells_kernel(ctx, B["cell_hash"], B["Y_wrapped"], _d3(d["inv_cell_size"]),
                                              _i3(d["num_cells"])))
    _check(L, ctx, L.ffcm_build_cell_list_kernel(ctx))
    _check(L, ctx, L.ffcm_gather_particles_kernel(ctx, B["Y_wrapped"], B["F_input"]))
    _check(L, ctx, L.ffcm_spread_forces_kernel(ctx, sig, Sig, h, ih, M, MG))
    for c in range(3):
        _check(L, ctx, L.ffcm_forward_transform(ctx, c))
    Mtot = int(np.prod(d["num_grid_points"]))
    inv_M = T(1) / T(Mtot)
    _check(L, ctx, L.ffcm_apply_inverse_stokes_kernel(ctx, float(d["mu"]), float(inv_M)))
    for c in range(3):
        _check(L, ctx, L.ffcm_inverse_transform(ctx, c))
    _check(L, ctx, L.ffcm_interpolate_velocities_kernel(ctx, B["V_output"], sig, Sig, h, ih, M, MG))
    _check(L, ctx, L.ffcm_correct_velocities_kernel(ctx, B["V_output"], _i3(d["num_cells"]), _d3(d["inv_cell_size"]),
                                                    _d3(d["L"]), sig, Sig, float(d["a"]), float(d["mu"]),
                                                    float(d["R_c"])))
    return _get(L, ctx, B["V_output"], (3, N), T)  # _retrieve_mobility_output!


@pytest.mark.parametrize("T,tol", [(np.float32, REL_L2_F32), (np.float64, 1e-11)])
def test_reference_call_sequence_matches_oracle(T, tol):
    from ForceCouplingMethod import FFCMConfig, _lib, mobility_
    from ForceCouplingMethod.config import derive_scalars, params_of
    from ForceCouplingMethod.suspensions import plain_workload

    L = _lib.lib()
    kw, Y, F = plain_workload(0).typed(T)
    d = derive_scalars(**kw)
    ctx = _assemble(L, d, T)
    try:
        # the fused entry points refuse to run before they know the scalars
        V = np.empty(Y.shape, T, order="F")
        assert L.ffcm_mobility(ctx, V.ctypes.data, Y.ctypes.data, F.ctypes.data, d["N"]) == _lib.FFCM_ERR_ARGUMENT
        assert "ffcm_set_scalars" in _lib.last_error(ctx)
        Vg = _generic_mobility(L, ctx, d, Y, F)
        Vo = orc.mobility(orc.OracleConfig(dtype=T, **kw), np.ascontiguousarray(Y.T), np.ascontiguousarray(F.T)).T
        e = rel_l2(Vg, Vo)
        print(f"reference call sequence ({T.__name__}): rel L2 vs CPU {e:.3e}")
        assert e <= tol
        # the neighbour map is the caller's, on the device
        nm = np.empty(13 * int(np.prod(d["num_cells"])), np.int32)
        _check(L, ctx, L.ffcm_get_buffer(ctx, _lib.BUFFER_IDS["neighbor_map"], nm.ctypes.data, nm.nbytes))
        assert np.array_equal(nm, neighbor_map(d["num_cells"]))
        # the fused mobility! method: every scalar at once, then one call
        p = params_of(d)
        _check(L, ctx, L.ffcm_set_scalars(ctx, ctypes.byref(p)))
        _check(L, ctx, L.ffcm_mobility(ctx, V.ctypes.data, Y.ctypes.data, F.ctypes.data, d["N"]))
        assert rel_l2(V, Vo) <= tol
        # and the ffcm_create context of the same config
        if T is np.float64:
            with pytest.warns(UserWarning):
                cfg = FFCMConfig(gpu_acceleration=True, **kw)
        else:
            cfg = FFCMConfig(gpu_acceleration=True, **kw)
        Vc = mobility_(np.empty(Y.shape, T, order="F"), cfg, Y, F)
        cfg.close()
        assert rel_l2(Vg, Vc) <= (1e-6 if T is np.float32 else 1e-13)
    finally:
        L.ffcm_destroy(ctx)


def test_forward_transforms_leave_the_unprojected_spectrum():
    # mul!(f̂_c, forward plan, f_c) alone: fluid_hat holds the r2c of each force component
    # (src/stokes_solve.jl:24-26), the projection not yet applied
    from ForceCouplingMethod import _lib
    from ForceCouplingMethod.config import derive_scalars

    T = np.float32
    L = _lib.lib()
    kw = dict(L=(T(12), T(16), T(20)), R_c=T(2), N=16, kernel_widths_ratio=T(2), num_grid_points=(24, 32, 40),
              M_G=6, viscosity=T(1))
    d = derive_scalars(**kw)
    ctx = _assemble(L, d, T)
    try:
        rng = np.random.default_rng(12)
        f = rng.normal(size=(3, 24, 32, 40)).astype(T)
        fa = np.ascontiguousarray(np.transpose(f, (0, 3, 2, 1)))  # (3, Mz, My, Mx) = column-major (3, Mx, My, Mz)
        _check(L, ctx, L.ffcm_set_buffer(ctx, _lib.BUFFER_IDS["force_density"], fa.ctypes.data, fa.nbytes))
        for c in range(3):
            _check(L, ctx, L.ffcm_forward_transform(ctx, c))
        sh = np.empty((3, 40, 32, 13), np.complex64)
        _check(L, ctx, L.ffcm_get_buffer(ctx, _lib.BUFFER_IDS["fluid_hat"], sh.ctypes.data, sh.nbytes))
        for c in range(3):
            ref = np.fft.rfftn(fa[c].astype(np.float64), axes=(0, 1, 2))  # real axis last: x
            assert rel_l2(sh[c], ref) <= 2e-6
    finally:
        L.ffcm_destroy(ctx)
