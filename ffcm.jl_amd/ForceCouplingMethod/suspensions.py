"""Seeded synthetic suspensions for the benchmark workloads and the parity tests.

The positions come from random sequential addition (RSA) of unit spheres in a cubic periodic
box (a candidate whose minimum-image distance to an accepted centre is below 2 is redrawn),
optionally with part of the particles drawn around Gaussian cluster centres (a clustered
particle whose cluster draws are rejected max_cluster_tries times is drawn uniformly; the
number of such fallbacks is returned); forces are i.i.d. N(0, 1) with the per-component mean
subtracted. Everything is generated in float64 by libffcm_suspension.so (project-written
xoshiro256** and counter-based streams; the cluster draws of one particle are tested in
parallel, with the same result for any thread count) and is identical on every machine;
single-precision runs use the rounded values.
"""
from __future__ import annotations

import ctypes
import functools
import math
import os
from dataclasses import dataclass, field

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(os.path.dirname(_PKG), "libffcm_suspension.so")
_lib = None


def _L():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} is not built (run __graft_entry__.build())")
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.ffcm_rsa_suspension.argtypes = [ctypes.c_int64, ctypes.c_double, ctypes.c_int, ctypes.c_int64,
                                             ctypes.c_double, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                                             ctypes.POINTER(ctypes.c_int64)]
        _lib.ffcm_normal_forces.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p]
        _lib.ffcm_uniform.argtypes = [ctypes.c_int64, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_void_p]
    return _lib


def rsa_positions(N, L, seed, n_clusters=0, n_clustered=0, cluster_std=8.0, max_cluster_tries=10000):
    """3xN float64 positions (Fortran order) and the number of cluster draws that fell back to
    a uniform draw."""
    Y = np.empty((3, N), dtype=np.float64, order="F")
    fb = ctypes.c_int64(0)
    rc = _L().ffcm_rsa_suspension(N, float(L), n_clusters, n_clustered, float(cluster_std), seed,
                                  max_cluster_tries, Y.ctypes.data, ctypes.byref(fb))
    if rc != 0:
        raise RuntimeError("random sequential addition could not place every particle (box too full)")
    return Y, int(fb.value)


def set_threads(n: int) -> None:
    """Threads that test cluster draws (the positions are the same for any count)."""
    _L().ffcm_suspension_set_threads(int(n))


def normal_forces(N, seed, n_rhs=1):
    """n_rhs sets of mean-subtracted N(0,1) forces: (3, N) for one set, else (n_rhs, 3, N)."""
    F = np.empty((n_rhs, N, 3), dtype=np.float64)
    _L().ffcm_normal_forces(N, n_rhs, seed, F.ctypes.data)
    F = np.transpose(F, (0, 2, 1))  # (n_rhs, 3, N), column n = particle n
    return np.asfortranarray(F[0]) if n_rhs == 1 else F


def uniform(n, seed, lo, hi):
    out = np.empty(n, dtype=np.float64)
    _L().ffcm_uniform(n, seed, float(lo), float(hi), out.ctypes.data)
    return out


def smooth5_at_least(n):
    """Smallest 2^a 3^b 5^c >= n."""
    m = max(1, int(n))
    while True:
        r = m
        for f in (2, 3, 5):
            while r % f == 0:
                r //= f
        if r == 1:
            return m
        m += 1


@dataclass
class Workload:
    name: str
    config: dict  # FFCMConfig keyword arguments (working precision applied by the caller)
    Y: np.ndarray  # 3xN float64
    F: np.ndarray  # 3xN float64
    info: dict = field(default_factory=dict)

    def typed(self, T):
        """Config kwargs with every real in precision T, and Y, F rounded to T."""
        kw = dict(self.config)
        kw["L"] = tuple(T(x) for x in kw["L"])
        for k in ("R_c", "kernel_widths_ratio", "viscosity"):
            kw[k] = T(kw[k])
        return kw, np.asfortranarray(self.Y.astype(T)), np.asfortranarray(self.F.astype(T))


@functools.lru_cache(maxsize=4)
def plain_workload(index: int) -> Workload:
    """The three plain suspensions of the benchmark, by index:
    0: N=512, L=32, 64^3 grid (RSA, uniform);
    1: N=65,536, L=140, 256^3 grid (RSA, phi ~ 0.10);
    2: N=1,048,576, L=280, 512^3 grid (RSA, phi ~ 0.20, half of the particles in 64 Gaussian
       clusters of std 8; a particle whose cluster draws are rejected 10,000 times is drawn
       uniformly, about 9 % of the clustered half).
    All: a=1, M_G=8, kernel_widths_ratio=2, R_c=4, mu=1; seed 0xFFC0 + index."""
    seed = 0xFFC0 + index
    if index == 0:
        N, L, M, ncl, nclus = 512, 32.0, 64, 0, 0
    elif index == 1:
        N, L, M, ncl, nclus = 65536, 140.0, 256, 0, 0
    elif index == 2:
        N, L, M, ncl, nclus = 1048576, 280.0, 512, 64, 1048576 // 2
    else:
        raise ValueError(index)
    Y, fb = rsa_positions(N, L, seed, n_clusters=ncl, n_clustered=nclus, cluster_std=8.0, max_cluster_tries=10000)
    F = normal_forces(N, seed ^ 0x5EED)
    cfg = dict(L=(L, L, L), R_c=4.0, N=N, kernel_widths_ratio=2.0, num_grid_points=(M, M, M), M_G=8,
               viscosity=1.0)
    phi = N * 4.0 / 3.0 * math.pi / L ** 3
    return Workload(f"plain{index}", cfg, Y, F, {"phi": phi, "cluster_fallbacks": fb, "seed": seed})


def ensemble_sizes(n_systems=64, seed=0xFFC3, phi=0.10, h=0.55):
    """(N_s, L_s, M_s) of the ensemble workload: N_s log-uniform in [1000, 50000], L_s sized to
    volume fraction phi, M_s the smallest 5-smooth size with L_s / M_s <= h."""
    u = uniform(n_systems, seed, math.log(1e3), math.log(5e4))
    out = []
    for x in u:
        N = int(round(math.exp(x)))
        L = (N * 4.0 * math.pi / 3.0 / phi) ** (1.0 / 3.0)
        M = smooth5_at_least(math.ceil(L / h))
        out.append((N, L, M))
    return out


@dataclass
class EnsembleSystem:
    config: dict  # FFCMConfig keyword arguments (float64 reals; cast by the caller)
    Y: np.ndarray  # 3 x N float64
    F: np.ndarray  # 3 x N x R float64 (Fortran order)

    def typed(self, T):
        kw = dict(self.config)
        kw["L"] = tuple(T(x) for x in kw["L"])
        for k in ("R_c", "kernel_widths_ratio", "viscosity"):
            kw[k] = T(kw[k])
        return kw, np.asfortranarray(self.Y, dtype=T), np.asfortranarray(self.F, dtype=T)


def ensemble_workload(n_systems=64, n_rhs=16, seed=0xFFC3, phi=0.10, h=0.55):
    """The batched ensemble: ``ensemble_sizes`` systems, each placed by random sequential
    addition (seed + 1 + k) with ``n_rhs`` mean-subtracted N(0,1) force sets (seed ^ 0x5EED + k).
    a=1, M_G=8, kernel_widths_ratio=2, R_c=4, mu=1."""
    out = []
    for k, (N, L, M) in enumerate(ensemble_sizes(n_systems, seed, phi, h)):
        Y, _ = rsa_positions(N, L, seed + 1 + k)
        F = normal_forces(N, (seed ^ 0x5EED) + k, n_rhs=n_rhs)
        if n_rhs == 1:
            F = F[:, :, None]
        else:
            F = np.transpose(F, (1, 2, 0))  # (3, N, R)
        cfg = dict(L=(L, L, L), R_c=4.0, N=N, kernel_widths_ratio=2.0, num_grid_points=(M, M, M), M_G=8,
                   viscosity=1.0)
        out.append(EnsembleSystem(cfg, Y, np.asfortranarray(F)))
    return out
