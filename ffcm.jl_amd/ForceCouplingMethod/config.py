"""``FFCMConfig``: validation, derived scalars and the HIP context that owns every buffer.

Host-side mirror of ``FFCMConfig{T}`` (src/config.jl:466-546). Validation follows
``_validate_config_parameters`` (src/config.jl:267-312) check for check and raises the same
``ArgumentError``s before any device is touched; the derived scalars follow
``_derive_scalars`` (src/config.jl:326-359) operation for operation in the working precision
T, so the hashing and stencil-anchor scalars handed to the kernels round exactly as Julia's.
The buffers (src/config.jl:24-125) live on the device inside the backend context
(``_assemble_gpu_buffers``, src/config.jl:461-464, :495-502); ``get_buffer`` / ``set_buffer``
are the ``Array(x)`` / ``copyto!`` a reference test applies to a config field.
"""
from __future__ import annotations

import ctypes
import warnings
from typing import Optional

import numpy as np

from . import _lib
from .errors import ArgumentError, DeviceError, DimensionMismatch, OutOfGPUMemoryError

_PLANNING = ("estimate", "measure", "patient")


def _infer_T(L, T):
    if T is not None:
        T = np.dtype(T).type
    else:
        kinds = {np.asarray(x).dtype for x in L}
        if len(kinds) != 1:
            raise ArgumentError("the elements of L must share one floating-point type")
        T = kinds.pop().type
    if T not in (np.float32, np.float64):
        raise ArgumentError(
            "the element type of L must be a concrete floating-point type to infer the working precision; "
            f"got {T}"
        )
    return T


def wrap_wavevectors(M: int, L, T):
    """``_wrap_wavevectors`` (src/config.jl:233-237): 2 pi m / L in FFT wrap order, index
    M/2+1 holding +M/2."""
    twopi = T(2) * T(np.pi)
    half_plus_one = M // 2 + 1
    out = np.empty(M, dtype=T)
    for j in range(1, M + 1):
        m = (j - 1) if j <= half_plus_one else (j - 1 - M)
        out[j - 1] = twopi * T(m) / L
    return out


def raise_for(code: int, ctx, what: str):
    if code == _lib.FFCM_OK:
        return
    msg = f"{what}: {_lib.last_error(ctx)}"
    if code == _lib.FFCM_ERR_ARGUMENT:
        raise ArgumentError(msg)
    if code == _lib.FFCM_ERR_DIMENSION:
        raise DimensionMismatch(msg)
    if code == _lib.FFCM_ERR_MEMORY:
        raise OutOfGPUMemoryError(msg)
    raise DeviceError(msg)


def derive_scalars(*, L, R_c, N, kernel_widths_ratio, num_grid_points, M_G, viscosity, a=None,
                   fft_planning="measure", fft_threads=1, T=None) -> dict:
    """Validation (``_validate_config_parameters``, src/config.jl:267-312, same order and
    messages) and derived scalars (``_derive_scalars``, src/config.jl:326-359) in precision T.
    Touches no device."""
    T = _infer_T(L, T)
    L = tuple(T(x) for x in L)
    if len(L) != 3:
        raise ArgumentError("L must have three components")
    R_c = T(R_c)
    a = T(1) if a is None else T(a)
    ratio = T(kernel_widths_ratio)
    mu = T(viscosity)
    M = tuple(int(m) for m in num_grid_points)
    planning = str(fft_planning).lstrip(":")
    if not R_c > 0:
        raise ArgumentError("R_c must be positive")
    if not all(x > 0 for x in L):
        raise ArgumentError("L components must be positive")
    if not R_c <= min(L) / T(2):
        raise ArgumentError(
            f"R_c must be at most half the smallest box length;got R_c = {R_c}, min(L)/2 = {min(L) / T(2)}"
        )
    if not N > 0:
        raise ArgumentError("N must be positive")
    if not a == T(1):
        raise ArgumentError(f"non-unit particle radius not yet implemented; only a = 1 is supported, got a = {a}")
    if not ratio >= T(1):
        raise ArgumentError(f"kernel widths ratio must be at least 1; got {ratio}")
    if not M_G >= 2:
        raise ArgumentError(f"M_G must be at least 2; got {M_G}")
    if not all(m >= 1 for m in M):
        raise ArgumentError("num_grid_points components must each be >= 1")
    if not M_G <= min(M):
        raise ArgumentError(
            f"M_G must not exceed the grid on any axis; got M_G = {M_G}, min(num_grid_points) = {min(M)}"
        )
    if not mu > 0:
        raise ArgumentError(f"viscosity must be positive; got {mu}")
    if planning not in _PLANNING:
        raise ArgumentError(f"fft_planning must be :estimate, :measure, or :patient; got :{planning}")
    if not fft_threads >= 1:
        raise ArgumentError(f"fft_threads must be at least 1; got {fft_threads}")
    num_cells = tuple(max(int(np.floor(L[i] / R_c)), 3) for i in range(3))
    cell_size = tuple(L[i] / T(num_cells[i]) for i in range(3))
    inv_cell_size = tuple(T(1) / cell_size[i] for i in range(3))
    sigma = a / np.sqrt(T(np.pi))
    Sigma = ratio * sigma
    h_axis = tuple(L[i] / T(M[i]) for i in range(3))
    rel = np.sqrt(np.finfo(T).eps)
    if not (abs(h_axis[1] - h_axis[0]) <= rel * h_axis[0] and abs(h_axis[2] - h_axis[0]) <= rel * h_axis[0]):
        raise ArgumentError(f"anisotropic grid spacing not supported (uniform h assumed); L_i/M_i = {h_axis}")
    h = h_axis[0]
    inv_h = T(1) / h
    Mh = M[0] // 2 + 1
    k_x = np.array([T(2) * T(np.pi) * T(i - 1) / L[0] for i in range(1, Mh + 1)], dtype=T)
    return dict(T=T, L=L, R_c=R_c, N=int(N), a=a, kernel_widths_ratio=ratio, num_grid_points=M, M_G=int(M_G),
                mu=mu, fft_planning=planning, fft_threads=int(fft_threads), num_cells=num_cells,
                cell_size=cell_size, inv_cell_size=inv_cell_size, sigma=sigma, Sigma=Sigma, h=h, inv_h=inv_h,
                k_x=k_x, k_y=wrap_wavevectors(M[1], L[1], T), k_z=wrap_wavevectors(M[2], L[2], T))


def params_of(d) -> "_lib.Params":
    """``ffcm_params`` from the derived scalars of ``derive_scalars`` (every scalar in T). The
    returned struct points into ``d``'s wavevector arrays, which must outlive the create call."""
    p = _lib.Params()
    p.dtype = _lib.FFCM_F32 if d["T"] is np.float32 else _lib.FFCM_F64
    p.N = d["N"]
    for i in range(3):
        p.M[i] = d["num_grid_points"][i]
        p.num_cells[i] = d["num_cells"][i]
        p.L[i] = float(d["L"][i])
        p.cell_size[i] = float(d["cell_size"][i])
        p.inv_cell_size[i] = float(d["inv_cell_size"][i])
    p.M_G = d["M_G"]
    p.R_c = float(d["R_c"])
    p.a = float(d["a"])
    p.sigma = float(d["sigma"])
    p.Sigma = float(d["Sigma"])
    p.h = float(d["h"])
    p.inv_h = float(d["inv_h"])
    p.mu = float(d["mu"])
    p.k_x = d["k_x"].ctypes.data
    p.k_y = d["k_y"].ctypes.data
    p.k_z = d["k_z"].ctypes.data
    return p


class FFCMConfig:
    """The compiled configuration of one FFCM operator (``FFCMConfig{T}``, src/config.jl:196-215).

    Keyword arguments are the reference's (src/config.jl:466-478). ``T`` is inferred from the
    element type of ``L`` (src/config.jl:539-546) unless given. Only ``gpu_acceleration=True``
    is served here: this package is the GPU backend; ``fft_planning`` and ``fft_threads`` are
    validated but have no effect on the device path.
    """

    def __init__(self, *, L, R_c, N, kernel_widths_ratio, num_grid_points, M_G, viscosity, a=None,
                 fft_planning="measure", fft_threads=1, gpu_acceleration=False, T=None, max_rhs=1):
        d = derive_scalars(L=L, R_c=R_c, N=N, kernel_widths_ratio=kernel_widths_ratio,
                           num_grid_points=num_grid_points, M_G=M_G, viscosity=viscosity, a=a,
                           fft_planning=fft_planning, fft_threads=fft_threads, T=T)
        self.__dict__.update(d)
        self.gpu_acceleration = bool(gpu_acceleration)
        self.max_rhs = int(max_rhs)  # right-hand sides per mobility_multi_ call (backend extension)
        self._ctx = None
        self._positions_owner = None  # the FFCMMobility whose cell list the context holds
        if not gpu_acceleration:
            raise ArgumentError(
                "this package is the gpu_acceleration backend of FFCMConfig; "
                "construct the config with gpu_acceleration=True"
            )
        if self.T is np.float64:
            warnings.warn("Float64 throughput is a fraction of Float32 on the GPU; consider Float32 for GPU runs")
        self._assemble_gpu_buffers()

    # _assemble_gpu_buffers (src/config.jl:461-464, :495-502)
    def _assemble_gpu_buffers(self):
        L = _lib.lib()
        p = params_of(self.__dict__)
        ctx = ctypes.c_void_p()
        rc = L.ffcm_create_multi(ctypes.byref(ctx), ctypes.byref(p), self.max_rhs)
        if rc != _lib.FFCM_OK:
            msg = _lib.last_error(None)
            if rc == _lib.FFCM_ERR_MEMORY:
                raise OutOfGPUMemoryError(msg)
            # a missing device is the reference's missing-backend ArgumentError
            raise ArgumentError(msg)
        self._ctx = ctx

    def close(self):
        if self._ctx is not None:
            _lib.lib().ffcm_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ctx(self):
        if self._ctx is None:
            raise ArgumentError("the configuration has been closed")
        return self._ctx

    def _check(self, rc, what):
        raise_for(rc, self._ctx, what)

    # ----------------------------------------------------------------- buffers
    def _shape(self, name):
        N = self.N
        Mx, My, Mz = self.num_grid_points
        if name in ("Y_input", "F_input", "V_output", "Y_wrapped", "Y_sorted", "F_sorted"):
            return (3, N), self.T
        if name in ("cell_hash", "original_index"):
            return (N,), np.int32
        if name in ("cell_start", "cell_end"):
            return (int(np.prod(self.num_cells)),), np.int32
        if name == "neighbor_map":
            return (13 * int(np.prod(self.num_cells)),), np.int32
        if name in ("force_density", "fluid_velocity"):
            return (3, Mx, My, Mz), self.T
        if name == "fluid_hat":
            return (3, Mx // 2 + 1, My, Mz), np.complex64 if self.T is np.float32 else np.complex128
        raise ArgumentError(f"unknown buffer {name}")

    def get_buffer(self, name: str) -> np.ndarray:
        """Device -> host copy of one config field in the reference layout: particle matrices
        as 3xN (Fortran order, column = particle); grid fields as (3, Mx, My, Mz) whose
        component slices are the reference's column-major component arrays."""
        shape, dt = self._shape(name)
        out = np.empty(shape, dtype=dt, order="F" if len(shape) == 2 else "C")
        if len(shape) == 4:
            out = np.empty((shape[0],) + shape[1:][::-1], dtype=dt)  # (3, Mz, My, Mx) C order
        nbytes = out.nbytes
        self._check(_lib.lib().ffcm_get_buffer(self.ctx, _lib.BUFFER_IDS[name], out.ctypes.data, nbytes), name)
        if len(shape) == 4:
            out = np.transpose(out, (0, 3, 2, 1))
        return out

    def set_buffer(self, name: str, value) -> None:
        """Host -> device copy into one config field (``copyto!(config.<field>, value)``)."""
        shape, dt = self._shape(name)
        arr = np.asarray(value, dtype=dt)
        if arr.shape != shape:
            raise DimensionMismatch(f"{name} must have shape {shape}; got {arr.shape}")
        if len(shape) == 2:
            arr = np.asfortranarray(arr)
        elif len(shape) == 4:
            arr = np.ascontiguousarray(np.transpose(arr, (0, 3, 2, 1)))
        else:
            arr = np.ascontiguousarray(arr)
        self._check(_lib.lib().ffcm_set_buffer(self.ctx, _lib.BUFFER_IDS[name], arr.ctypes.data, arr.nbytes), name)

    # ----------------------------------------------------------------- profiling
    def info(self) -> dict:
        tile = (ctypes.c_int32 * 3)()
        nt = ctypes.c_int32()
        fs = (ctypes.c_int32 * 12)()
        ik = ctypes.c_int32()
        self._check(_lib.lib().ffcm_get_info(self.ctx, tile, ctypes.byref(nt), fs, ctypes.byref(ik)), "info")
        names = ("x_forward", "y", "z_project", "x_inverse")
        return {
            "tile": tuple(tile),
            "n_tiles": nt.value,
            "interp": ("wave", "column", "dma")[ik.value],
            "fft": {names[i]: {"width": fs[3 * i], "threads": fs[3 * i + 1], "lds_bytes": fs[3 * i + 2]}
                    for i in range(4)},
        }

    def set_timing(self, enable: bool) -> None:
        self._check(_lib.lib().ffcm_set_timing(self.ctx, 1 if enable else 0), "set_timing")

    def set_graph(self, enable: bool) -> None:
        """Device-resident applies replayed from a captured hipGraph while their pointers stay
        the same (the first apply runs as launched, the second is captured)."""
        self._check(_lib.lib().ffcm_set_graph(self.ctx, 1 if enable else 0), "set_graph")

    def timing(self) -> dict:
        ms = (ctypes.c_double * len(_lib.STAGES))()
        calls = ctypes.c_int64()
        self._check(_lib.lib().ffcm_get_timing(self.ctx, ms, len(_lib.STAGES), ctypes.byref(calls)), "timing")
        return {"calls": calls.value, "ms": dict(zip(_lib.STAGES, list(ms)))}

    def stream(self) -> int:
        s = ctypes.c_void_p()
        self._check(_lib.lib().ffcm_get_stream(self.ctx, ctypes.byref(s)), "stream")
        return s.value or 0

    def synchronize(self) -> None:
        self._check(_lib.lib().ffcm_synchronize(self.ctx), "synchronize")

    def __repr__(self):
        return (f"FFCMConfig{{{self.T.__name__}}}(N={self.N}, L={self.L}, num_grid_points={self.num_grid_points}, "
                f"M_G={self.M_G}, R_c={self.R_c}, num_cells={self.num_cells}, gpu_acceleration=True)")
