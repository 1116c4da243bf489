"""ctypes front end of the CPU restatement in ``ffcm_oracle.cpp``.

TEST INFRASTRUCTURE ONLY: the parity oracle for the HIP backend. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import this module, and
only as the checker. The product (``ffcm.jl_amd``) never imports it.

``OracleConfig`` restates ``FFCMConfig{T}``'s validation and derived scalars
(src/config.jl:267-359, :233-237) in numpy scalars of the working precision, so every
rounding matches Julia's (IEEE single/double with correctly rounded + - * / sqrt).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("dtype", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("N", ctypes.c_int64),
        ("M", ctypes.c_int32 * 3),
        ("M_G", ctypes.c_int32),
        ("num_cells", ctypes.c_int32 * 3),
        ("pad1", ctypes.c_int32),
        ("L", ctypes.c_double * 3),
        ("R_c", ctypes.c_double),
        ("a", ctypes.c_double),
        ("sigma", ctypes.c_double),
        ("Sigma", ctypes.c_double),
        ("h", ctypes.c_double),
        ("inv_h", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("cell_size", ctypes.c_double * 3),
        ("inv_cell_size", ctypes.c_double * 3),
        ("k_x", ctypes.c_void_p),
        ("k_y", ctypes.c_void_p),
        ("k_z", ctypes.c_void_p),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run __graft_entry__.build())")
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def set_threads(fft_threads: int = 1, particle_threads: int = 1) -> None:
    """Threads of the transforms and of the spread / interpolation / correction loops (results
    are bit-identical for any counts; 1 and 1 restate the reference's serial loops with
    fft_threads = 1)."""
    lib().oracle_set_threads(int(fft_threads), int(particle_threads))


def _wrap_wavevectors(M: int, L, T):
    """src/config.jl:233-237: k = 2π m / L in FFTW wrap order (index M/2+1 is +M/2)."""
    twopi = T(2) * T(np.pi)
    half_plus_one = M // 2 + 1
    out = np.empty(M, dtype=T)
    for j in range(1, M + 1):
        m = (j - 1) if j <= half_plus_one else (j - 1 - M)
        out[j - 1] = twopi * T(m) / L
    return out


@dataclass
class OracleConfig:
    """Scalars of ``FFCMConfig{T}`` (src/config.jl:466-537) for the CPU restatement."""

    L: tuple
    R_c: float
    N: int
    kernel_widths_ratio: float
    num_grid_points: tuple
    M_G: int
    viscosity: float
    a: float = 1.0
    dtype: type = np.float32

    def __post_init__(self):
        T = np.dtype(self.dtype).type
        self.T = T
        L = tuple(T(x) for x in self.L)
        R_c = T(self.R_c)
        M = tuple(int(m) for m in self.num_grid_points)
        # _derive_scalars (src/config.jl:326-359)
        self.num_cells = tuple(max(int(np.floor(L[i] / R_c)), 3) for i in range(3))
        self.cell_size = tuple(L[i] / T(self.num_cells[i]) for i in range(3))
        self.inv_cell_size = tuple(T(1) / self.cell_size[i] for i in range(3))
        self.a_T = T(self.a)
        self.sigma = self.a_T / np.sqrt(T(np.pi))
        self.Sigma = T(self.kernel_widths_ratio) * self.sigma
        h_axis = tuple(L[i] / T(M[i]) for i in range(3))
        rel = np.sqrt(np.finfo(T).eps)
        if not (abs(h_axis[1] - h_axis[0]) <= rel * h_axis[0] and abs(h_axis[2] - h_axis[0]) <= rel * h_axis[0]):
            raise ValueError("anisotropic grid spacing")
        self.h = h_axis[0]
        self.inv_h = T(1) / self.h
        self.mu = T(self.viscosity)
        Mh = M[0] // 2 + 1
        self.k_x = np.array([T(2) * T(np.pi) * T(i - 1) / L[0] for i in range(1, Mh + 1)], dtype=T)
        self.k_y = _wrap_wavevectors(M[1], L[1], T)
        self.k_z = _wrap_wavevectors(M[2], L[2], T)
        self.L_T = L
        self.R_c_T = R_c
        self.M = M

    @property
    def suffix(self):
        return "f32" if self.T is np.float32 else "f64"

    @property
    def total_cells(self):
        return self.num_cells[0] * self.num_cells[1] * self.num_cells[2]

    def params(self) -> OracleParams:
        p = OracleParams()
        p.dtype = 0 if self.T is np.float32 else 1
        p.N = int(self.N)
        for i in range(3):
            p.M[i] = self.M[i]
            p.num_cells[i] = self.num_cells[i]
            p.L[i] = float(self.L_T[i])
            p.cell_size[i] = float(self.cell_size[i])
            p.inv_cell_size[i] = float(self.inv_cell_size[i])
        p.M_G = int(self.M_G)
        p.R_c = float(self.R_c_T)
        p.a = float(self.a_T)
        p.sigma = float(self.sigma)
        p.Sigma = float(self.Sigma)
        p.h = float(self.h)
        p.inv_h = float(self.inv_h)
        p.mu = float(self.mu)
        p.k_x = self.k_x.ctypes.data
        p.k_y = self.k_y.ctypes.data
        p.k_z = self.k_z.ctypes.data
        self._keep = p  # the struct points into self.k_*; keep both alive together
        return p


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _col(a, T):
    """A 3xN column-major matrix as the (N, 3) C-order array the library reads."""
    a = np.ascontiguousarray(np.asarray(a, dtype=T))
    assert a.ndim == 2 and a.shape[1] == 3
    return a


def wrap_positions(cfg: OracleConfig, Y):
    Y = _col(Y, cfg.T)
    out = np.empty_like(Y)
    getattr(lib(), f"oracle_wrap_{cfg.suffix}")(ctypes.byref(cfg.params()), _ptr(Y), _ptr(out))
    return out


def assign_cells(cfg: OracleConfig, Yw):
    Yw = _col(Yw, cfg.T)
    h = np.empty(cfg.N, dtype=np.int32)
    getattr(lib(), f"oracle_assign_cells_{cfg.suffix}")(ctypes.byref(cfg.params()), _ptr(Yw), _ptr(h))
    return h


def build_cell_list(hash_, total):
    hash_ = np.ascontiguousarray(hash_, dtype=np.int32)
    N = hash_.shape[0]
    orig = np.empty(N, dtype=np.int32)
    cs = np.empty(total, dtype=np.int32)
    ce = np.empty(total, dtype=np.int32)
    lib().oracle_build_cell_list(ctypes.c_int64(N), ctypes.c_int64(total), _ptr(hash_), _ptr(orig), _ptr(cs), _ptr(ce))
    return orig, cs, ce


def gather(cfg: OracleConfig, orig, Y, F):
    Y = _col(Y, cfg.T)
    F = _col(F, cfg.T)
    orig = np.ascontiguousarray(orig, dtype=np.int32)
    Ys = np.empty_like(Y)
    Fs = np.empty_like(F)
    getattr(lib(), f"oracle_gather_{cfg.suffix}")(ctypes.byref(cfg.params()), _ptr(orig), _ptr(Y), _ptr(F), _ptr(Ys), _ptr(Fs))
    return Ys, Fs


def spread(cfg: OracleConfig, Ys, Fs):
    """Returns the force density as a (3, Mz, My, Mx) array (x fastest in memory)."""
    Ys = _col(Ys, cfg.T)
    Fs = _col(Fs, cfg.T)
    Mx, My, Mz = cfg.M
    grid = np.empty((3, Mz, My, Mx), dtype=cfg.T)
    getattr(lib(), f"oracle_spread_{cfg.suffix}")(ctypes.byref(cfg.params()), _ptr(Ys), _ptr(Fs), _ptr(grid))
    return grid


def rfft3(M, field, T):
    """Unnormalised r2c of one (Mz, My, Mx) real field -> (Mz, My, Mx//2+1) complex."""
    field = np.ascontiguousarray(field, dtype=T)
    Mx, My, Mz = M
    C = np.complex64 if T is np.float32 else np.complex128
    out = np.empty((Mz, My, Mx // 2 + 1), dtype=C)
    Marr = np.array(M, dtype=np.int32)
    sfx = "f32" if T is np.float32 else "f64"
    getattr(lib(), f"oracle_rfft3_{sfx}")(_ptr(Marr), _ptr(field), _ptr(out))
    return out


def brfft3(M, spec, T):
    """Unnormalised c2r of (Mz, My, Mx//2+1) complex -> (Mz, My, Mx) real (FFTW brfft)."""
    C = np.complex64 if T is np.float32 else np.complex128
    spec = np.array(spec, dtype=C, copy=True, order="C")
    Mx, My, Mz = M
    out = np.empty((Mz, My, Mx), dtype=T)
    Marr = np.array(M, dtype=np.int32)
    sfx = "f32" if T is np.float32 else "f64"
    getattr(lib(), f"oracle_brfft3_{sfx}")(_ptr(Marr), _ptr(spec), _ptr(out))
    return out


def stokes(cfg: OracleConfig, fgrid):
    """Returns (velocity grid (3,Mz,My,Mx), projected spectrum (3,Mz,My,Mx//2+1))."""
    fgrid = np.ascontiguousarray(fgrid, dtype=cfg.T)
    Mx, My, Mz = cfg.M
    C = np.complex64 if cfg.T is np.float32 else np.complex128
    u = np.empty((3, Mz, My, Mx), dtype=cfg.T)
    fh = np.empty((3, Mz, My, Mx // 2 + 1), dtype=C)
    getattr(lib(), f"oracle_stokes_{cfg.suffix}")(ctypes.byref(cfg.params()), _ptr(fgrid), _ptr(u), _ptr(fh))
    return u, fh


def interpolate(cfg: OracleConfig, ugrid, Ys, orig):
    ugrid = np.ascontiguousarray(ugrid, dtype=cfg.T)
    Ys = _col(Ys, cfg.T)
    orig = np.ascontiguousarray(orig, dtype=np.int32)
    V = np.zeros((cfg.N, 3), dtype=cfg.T)
    getattr(lib(), f"oracle_interp_{cfg.suffix}")(ctypes.byref(cfg.params()), _ptr(ugrid), _ptr(Ys), _ptr(orig), _ptr(V))
    return V


def correct(cfg: OracleConfig, Ys, Fs, cs, ce, orig, V):
    """V += correction, in place; V is (N, 3)."""
    Ys = _col(Ys, cfg.T)
    Fs = _col(Fs, cfg.T)
    assert V.dtype == cfg.T and V.flags.c_contiguous
    args = [np.ascontiguousarray(x, dtype=np.int32) for x in (cs, ce, orig)]
    getattr(lib(), f"oracle_correct_{cfg.suffix}")(
        ctypes.byref(cfg.params()), _ptr(Ys), _ptr(Fs), _ptr(args[0]), _ptr(args[1]), _ptr(args[2]), _ptr(V)
    )
    return V


def mobility(cfg: OracleConfig, Y, F):
    """Full mobility! on the CPU restatement; Y, F, result are (N, 3) arrays."""
    Y = _col(Y, cfg.T)
    F = _col(F, cfg.T)
    V = np.zeros((cfg.N, 3), dtype=cfg.T)
    rc = getattr(lib(), f"oracle_mobility_{cfg.suffix}")(ctypes.byref(cfg.params()), _ptr(Y), _ptr(F), _ptr(V))
    assert rc == 0
    return V


def mobility_vjp(cfg: OracleConfig, Y, F, G):
    """Pullback of mobility! for L = <G, M(Y) F>: returns (dF, dY), each (N, 3)."""
    Y = _col(Y, cfg.T)
    F = _col(F, cfg.T)
    G = _col(G, cfg.T)
    dF = np.zeros((cfg.N, 3), dtype=cfg.T)
    dY = np.zeros((cfg.N, 3), dtype=cfg.T)
    rc = getattr(lib(), f"oracle_mobility_vjp_{cfg.suffix}")(
        ctypes.byref(cfg.params()), _ptr(Y), _ptr(F), _ptr(G), _ptr(dF), _ptr(dY)
    )
    assert rc == 0
    return dF, dY


def correction_scalars_deriv(r, sigma, Sigma, mu, T):
    """(A'(r), B'(r)) of the pair-correction scalars."""
    sfx = "f32" if T is np.float32 else "f64"
    ct = ctypes.c_float if T is np.float32 else ctypes.c_double
    dA = ct()
    dB = ct()
    f = getattr(lib(), f"oracle_correction_scalars_deriv_{sfx}")
    f.argtypes = [ct, ct, ct, ct, ctypes.POINTER(ct), ctypes.POINTER(ct)]
    f(ct(r), ct(sigma), ct(Sigma), ct(mu), ctypes.byref(dA), ctypes.byref(dB))
    return T(dA.value), T(dB.value)


def correction_scalars(r, sigma, Sigma, mu, T):
    sfx = "f32" if T is np.float32 else "f64"
    ct = ctypes.c_float if T is np.float32 else ctypes.c_double
    A = ct()
    B = ct()
    f = getattr(lib(), f"oracle_correction_scalars_{sfx}")
    f.argtypes = [ct, ct, ct, ct, ctypes.POINTER(ct), ctypes.POINTER(ct)]
    f(ct(r), ct(sigma), ct(Sigma), ct(mu), ctypes.byref(A), ctypes.byref(B))
    return T(A.value), T(B.value)


def self_correction(sigma, Sigma, a, mu, T):
    sfx = "f32" if T is np.float32 else "f64"
    ct = ctypes.c_float if T is np.float32 else ctypes.c_double
    f = getattr(lib(), f"oracle_self_correction_{sfx}")
    f.argtypes = [ct, ct, ct, ct]
    f.restype = ct
    return T(f(ct(sigma), ct(Sigma), ct(a), ct(mu)))


def kernel_coefficients(sigma, Sigma, T):
    sfx = "f32" if T is np.float32 else "f64"
    ct = ctypes.c_float if T is np.float32 else ctypes.c_double
    out = np.empty(4, dtype=T)
    f = getattr(lib(), f"oracle_kernel_coefficients_{sfx}")
    f.argtypes = [ct, ct, ctypes.c_void_p]
    f(ct(sigma), ct(Sigma), out.ctypes.data)
    return tuple(T(x) for x in out)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
