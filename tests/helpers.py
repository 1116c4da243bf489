"""Helpers shared by the parity tests: paired GPU / oracle configurations and tolerances.

Tolerances follow the reference suite (test/test_utilities.jl:6-9, test/README.md:25-31):
relative sqrt(eps(T)) with a near-zero absolute floor; end-to-end velocities must agree with
the CPU Float32 restatement within relative L2 <= 1e-5 (spreading and FFT summation order
differ).
"""
import numpy as np

from oracle import oracle as orc

REL_L2_F32 = 1e-5
REL_L2_F64 = 1e-10


def rtol_of(T):
    return float(np.sqrt(np.finfo(T).eps))


def atol_of(T):
    return 1e-6 if T is np.float32 else 1e-10


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) else np.float64)
    b = np.asarray(b, dtype=np.complex128 if np.iscomplexobj(b) else np.float64)
    nb = np.linalg.norm(b.ravel())
    d = np.linalg.norm((a - b).ravel())
    return float(d / nb) if nb > 0 else float(d)


def kwargs(T, L=(8.0, 8.0, 8.0), R_c=1.0, N=16, ratio=2.0, M=(16, 16, 16), M_G=8, mu=1.0):
    return dict(L=tuple(T(x) for x in L), R_c=T(R_c), N=N, kernel_widths_ratio=T(ratio), num_grid_points=M,
                M_G=M_G, viscosity=T(mu))


def pair(T, **kw):
    """(GPU FFCMConfig, OracleConfig) built from the same keyword arguments."""
    from ForceCouplingMethod import FFCMConfig

    k = kwargs(T, **kw)
    gpu = FFCMConfig(gpu_acceleration=True, **k)
    ora = orc.OracleConfig(dtype=T, **k)
    return gpu, ora


def oracle_rows(x):
    """3xN (reference layout) -> (N, 3) C-order rows the oracle reads."""
    return np.ascontiguousarray(np.asarray(x).T)


def oracle_grid(g):
    """Oracle (3, Mz, My, Mx) grid -> (3, Mx, My, Mz) reference layout."""
    return np.transpose(g, (0, 3, 2, 1))


def neighbor_map(nc):
    """_build_neighbor_map (src/cell_list.jl:258-289): the 13 forward half-shell neighbour cells
    of every cell, 0-based, periodic, cell index x fastest (restated in numpy)."""
    off = [(1, 0, 0), (1, 1, 0), (0, 1, 0), (-1, 1, 0), (1, 0, -1), (1, 1, -1), (0, 1, -1), (-1, 1, -1),
           (1, 0, 1), (1, 1, 1), (0, 1, 1), (-1, 1, 1), (0, 0, 1)]
    mx, my, mz = nc
    cz, cy, cx = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    out = np.empty((mz, my, mx, 13), dtype=np.int32)
    for k, (dx, dy, dz) in enumerate(off):
        out[..., k] = (cx + dx) % mx + ((cy + dy) % my + ((cz + dz) % mz) * my) * mx
    return out.reshape(-1)
