"""CPU checks of the drop-in boundary: the C-ABI library loads and exports every symbol
include/ffcm_hip.h declares, and parameter validation raises the reference's errors before any
device is touched (test/accuracy/test_cell_geometry.jl:52-82, test_fcm_grid.jl:23-108,
test/cuda/test_gpu_backend_selection.jl:13-33)."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import PKG, ROOT
from oracle import oracle as orc

HEADER = os.path.join(ROOT, "include", "ffcm_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(ffcm_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_symbols()
    for must in ("ffcm_create", "ffcm_destroy", "ffcm_mobility", "ffcm_mobility_device", "ffcm_spread_forces",
                 "ffcm_stokes_solve", "ffcm_interpolate_velocities", "ffcm_correct_velocities"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from ForceCouplingMethod import _lib

    L = _lib.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name
    assert set(_lib.EXPORTS) == set(declared_symbols())


def test_params_struct_layout_matches_header():
    from ForceCouplingMethod import _lib

    # int32 dtype, pad, int64 N, int32 M[3], M_G, num_cells[3], pad, then 16 doubles and 3 pointers
    assert ctypes.sizeof(_lib.Params) == 4 + 4 + 8 + 12 + 4 + 12 + 4 + 8 * 16 + 8 * 3
    assert _lib.Params.L.offset == 48


def _params(**over):
    from ForceCouplingMethod import _lib

    p = _lib.Params()
    p.dtype = 0
    p.N = 8
    for i in range(3):
        p.M[i] = 16
        p.num_cells[i] = 8
        p.L[i] = 8.0
        p.cell_size[i] = 1.0
        p.inv_cell_size[i] = 1.0
    p.M_G = 8
    p.R_c = 1.0
    p.a = 1.0
    p.sigma = 1 / np.sqrt(np.pi)
    p.Sigma = 2 / np.sqrt(np.pi)
    p.h = 0.5
    p.inv_h = 2.0
    p.mu = 1.0
    k = np.zeros(16, dtype=np.float32)
    p.k_x = p.k_y = p.k_z = k.ctypes.data
    for key, v in over.items():
        if key in ("M", "L", "num_cells"):
            for i in range(3):
                getattr(p, key)[i] = v[i]
        else:
            setattr(p, key, v)
    return p, k


@pytest.mark.parametrize("over,needle", [
    (dict(R_c=0.0), "R_c must be positive"),
    (dict(L=(8.0, -1.0, 8.0)), "L components"),
    (dict(R_c=5.0), "half the smallest box"),
    (dict(N=0), "N must be positive"),
    (dict(a=2.0), "non-unit particle radius"),
    (dict(Sigma=0.1), "kernel widths ratio"),
    (dict(M_G=1), "M_G must be at least 2"),
    (dict(M_G=17), "M_G must not exceed"),
    (dict(mu=0.0), "viscosity"),
    (dict(M=(16, 17, 16)), "anisotropic"),
    (dict(M=(17, 17, 17), L=(8.5, 8.5, 8.5)), "primes <= 13"),
    (dict(dtype=7), "dtype"),
    (dict(N=2 ** 26 + 1), "N above 2^26"),
])
def test_create_rejects_invalid_parameters_before_touching_a_device(over, needle):
    from ForceCouplingMethod import _lib

    L = _lib.lib()
    p, keep = _params(**over)
    ctx = ctypes.c_void_p()
    rc = L.ffcm_create(ctypes.byref(ctx), ctypes.byref(p))
    assert rc == _lib.FFCM_ERR_ARGUMENT
    assert needle in _lib.last_error(None)
    assert not ctx.value


def test_config_validation_raises_argument_error():
    # test/accuracy/test_cell_geometry.jl:52-82, test_fcm_grid.jl:23-108
    from ForceCouplingMethod import ArgumentError, FFCMConfig

    T = np.float32
    base = dict(L=(T(8), T(8), T(8)), R_c=T(1), N=4, kernel_widths_ratio=T(2), num_grid_points=(16, 16, 16), M_G=8,
                viscosity=T(1), gpu_acceleration=True)
    bad = [dict(R_c=T(0)), dict(R_c=T(5)), dict(N=0), dict(a=T(2)), dict(kernel_widths_ratio=T(0.5)), dict(M_G=1),
           dict(M_G=32), dict(viscosity=T(-1)), dict(num_grid_points=(16, 8, 16)), dict(fft_planning="fast"),
           dict(fft_threads=0), dict(L=(T(8), T(0), T(8)))]
    for b in bad:
        with pytest.raises(ArgumentError):
            FFCMConfig(**{**base, **b})
    with pytest.raises(ArgumentError):
        FFCMConfig(**{**base, "L": (8, 8, 8)})  # integer element type: no working precision


def test_derived_scalars_match_the_oracle_restatement():
    # both restate _derive_scalars (src/config.jl:326-359) in T; every scalar must be identical
    from ForceCouplingMethod.config import derive_scalars

    for T in (np.float32, np.float64):
        for L, M, Rc in (((32.0,) * 3, (64,) * 3, 4.0), ((140.0,) * 3, (256,) * 3, 4.0), ((4.0, 6.0, 8.0), (8, 12, 16), 1.0)):
            kw = dict(L=tuple(T(x) for x in L), R_c=T(Rc), N=10, kernel_widths_ratio=T(2), num_grid_points=M, M_G=8,
                      viscosity=T(1))
            d = derive_scalars(**kw)
            o = orc.OracleConfig(dtype=T, **kw)
            assert d["num_cells"] == o.num_cells
            for key in ("cell_size", "inv_cell_size"):
                assert all(a == b for a, b in zip(d[key], getattr(o, key)))
            for key in ("sigma", "Sigma", "h", "inv_h"):
                assert d[key] == getattr(o, key), key
            for key in ("k_x", "k_y", "k_z"):
                assert np.array_equal(d[key], getattr(o, key))


def test_missing_device_is_an_argument_error():
    # test/cuda/test_gpu_backend_selection.jl:13-33: no backend -> ArgumentError (subprocess)
    code = (
        "import sys, numpy as np\n"
        f"sys.path.insert(0, {PKG!r})\n"
        "from ForceCouplingMethod import FFCMConfig, ArgumentError\n"
        "T = np.float32\n"
        "try:\n"
        "    FFCMConfig(L=(T(8),)*3, R_c=T(1), N=4, kernel_widths_ratio=T(2), num_grid_points=(16,)*3, M_G=8,\n"
        "               viscosity=T(1), gpu_acceleration=True)\n"
        "except ArgumentError as e:\n"
        "    print('ArgumentError', e); sys.exit(42)\n"
        "sys.exit(0)\n"
    )
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 42, r.stdout + r.stderr


def test_cpu_path_is_not_served():
    from ForceCouplingMethod import ArgumentError, FFCMConfig

    T = np.float32
    with pytest.raises(ArgumentError, match="gpu_acceleration"):
        FFCMConfig(L=(T(8),) * 3, R_c=T(1), N=4, kernel_widths_ratio=T(2), num_grid_points=(16,) * 3, M_G=8,
                   viscosity=T(1))


def test_suspension_generator_is_deterministic_and_valid():
    from ForceCouplingMethod import suspensions as s

    Y1, _ = s.rsa_positions(2000, 30.0, 7)
    Y2, _ = s.rsa_positions(2000, 30.0, 7)
    assert np.array_equal(Y1, Y2)
    assert Y1.min() >= 0 and Y1.max() < 30.0
    # no pair closer than 2 (minimum image), brute force
    d = Y1[:, :, None] - Y1[:, None, :]
    d -= 30.0 * np.round(d / 30.0)
    r2 = (d ** 2).sum(axis=0) + np.eye(2000) * 100
    assert r2.min() >= 4.0
    F = s.normal_forces(2000, 3)
    assert F.shape == (3, 2000)
    np.testing.assert_allclose(F.mean(axis=1), 0, atol=1e-12)
    sizes = s.ensemble_sizes()
    assert len(sizes) == 64 and all(1000 <= n <= 50000 for n, _, _ in sizes)
    for n, L, M in sizes:
        assert L / M <= 0.55 and s.smooth5_at_least(M) == M


def test_clustered_suspension_is_thread_count_independent_and_valid():
    # the clustered draws of one particle are tested in parallel; the first free draw in draw
    # order must win for any thread count, and no accepted pair may overlap
    from ForceCouplingMethod import suspensions as s

    args = (3000, 36.0, 11)
    kw = dict(n_clusters=4, n_clustered=1500, cluster_std=4.0, max_cluster_tries=10000)
    try:
        s.set_threads(1)
        Y1, fb1 = s.rsa_positions(*args, **kw)
        s.set_threads(3)
        Y3, fb3 = s.rsa_positions(*args, **kw)
    finally:
        s.set_threads(os.cpu_count() or 1)
    assert np.array_equal(Y1, Y3) and fb1 == fb3
    assert 0 < fb1 < 1500  # the clusters jam: some fall back, most are placed in the clusters
    d = Y1[:, :, None] - Y1[:, None, :]
    d -= 36.0 * np.round(d / 36.0)
    r2 = (d ** 2).sum(axis=0) + np.eye(3000) * 100
    assert r2.min() >= 4.0
    # fewer tries per clustered particle: more fallbacks
    kw_few = dict(kw, max_cluster_tries=1)
    _, fb_few = s.rsa_positions(*args, **kw_few)
    assert fb_few > fb1


def test_oracle_is_bit_identical_for_any_thread_count():
    from oracle import oracle as orc
    from ForceCouplingMethod.suspensions import plain_workload

    kw, Y, F = plain_workload(0).typed(np.float32)
    cfg = orc.OracleConfig(dtype=np.float32, **kw)
    Yc, Fc = np.ascontiguousarray(Y.T), np.ascontiguousarray(F.T)
    try:
        orc.set_threads(1, 1)
        V1 = orc.mobility(cfg, Yc, Fc)
        orc.set_threads(3, 5)
        V3 = orc.mobility(cfg, Yc, Fc)
    finally:
        orc.set_threads(1, 1)
    assert np.array_equal(V1, V3)


def _seam_args(T=np.float32):
    from helpers import neighbor_map
    from ForceCouplingMethod.config import derive_scalars

    d = derive_scalars(L=(T(8), T(8), T(8)), R_c=T(1), N=8, kernel_widths_ratio=T(2), num_grid_points=(16, 16, 16),
                       M_G=8, viscosity=T(1))
    return d, neighbor_map(d["num_cells"])


def test_assemble_rejects_a_map_of_no_cell_grid_before_touching_a_device():
    # ffcm_assemble takes _assemble_gpu_buffers' nine arguments (src/config.jl:495-502) and
    # reads the per-axis cell counts from the half-shell map
    from ForceCouplingMethod import _lib

    L = _lib.lib()
    d, nm = _seam_args()
    C = int(np.prod(d["num_cells"]))
    M = (ctypes.c_int32 * 3)(16, 16, 16)
    ip = ctypes.POINTER(ctypes.c_int32)
    bad = nm.copy()
    bad[5] += 1
    ctx = ctypes.c_void_p()
    rc = L.ffcm_assemble(ctypes.byref(ctx), 0, 8, M, 8, C, d["k_x"].ctypes.data, d["k_y"].ctypes.data,
                         d["k_z"].ctypes.data, bad.ctypes.data_as(ip))
    assert rc == _lib.FFCM_ERR_ARGUMENT and not ctx.value
    assert "half-shell map" in _lib.last_error(None)
    rc = L.ffcm_assemble(ctypes.byref(ctx), 0, 8, M, 8, C - 1, d["k_x"].ctypes.data, d["k_y"].ctypes.data,
                         d["k_z"].ctypes.data, nm.ctypes.data_as(ip))
    assert rc == _lib.FFCM_ERR_ARGUMENT and not ctx.value
