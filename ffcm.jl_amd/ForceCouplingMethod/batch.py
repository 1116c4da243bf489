"""Batched ensemble of independent suspensions (no reference counterpart).

Each system k has its own ``FFCMConfig`` parameters and R right-hand sides sharing its
positions; ``V_k[:, :, r]`` equals ``mobility!`` on that system with ``F_k[:, :, r]``
(src/mobility.jl:23-45). The whole ensemble apply is one captured hipGraph in the HIP
backend (``ffcm_batch_*`` in include/ffcm_hip.h): one grouped launch per pipeline step over the
systems, or each system's own launches, on 1..4 stream lanes, whichever is fastest on the first
apply.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .config import derive_scalars, params_of
from .errors import ArgumentError, DeviceError, DimensionMismatch, OutOfGPUMemoryError


def _raise(code, batch, what):
    if code == _lib.FFCM_OK:
        return
    msg = _lib.lib().ffcm_batch_last_error(batch)
    msg = f"{what}: {msg.decode() if msg else ''}"
    if code == _lib.FFCM_ERR_ARGUMENT:
        raise ArgumentError(msg)
    if code == _lib.FFCM_ERR_DIMENSION:
        raise DimensionMismatch(msg)
    if code == _lib.FFCM_ERR_MEMORY:
        raise OutOfGPUMemoryError(msg)
    raise DeviceError(msg)


class FFCMBatch:
    """``systems``: a list of ``FFCMConfig`` keyword dictionaries (one working precision for
    all); ``n_rhs``: right-hand sides per system."""

    def __init__(self, systems, n_rhs: int = 1, T=None):
        if not systems:
            raise ArgumentError("a batch needs at least one system")
        self.derived = []
        for s in systems:
            kw = {k: v for k, v in dict(s).items() if k not in ("gpu_acceleration", "max_rhs")}
            if T is not None:
                kw["T"] = T
            self.derived.append(derive_scalars(**kw))
        Ts = {d["T"] for d in self.derived}
        if len(Ts) != 1:
            raise ArgumentError("every system of a batch must use the same element type")
        self.T = Ts.pop()
        self.n_rhs = int(n_rhs)
        self.N = [d["N"] for d in self.derived]
        self.K = len(systems)
        arr = (_lib.Params * self.K)(*[params_of(d) for d in self.derived])
        b = ctypes.c_void_p()
        rc = _lib.lib().ffcm_batch_create(ctypes.byref(b), self.K, arr, self.n_rhs)
        if rc != _lib.FFCM_OK:
            _raise(rc, None, "FFCMBatch")
        self._b = b

    def close(self):
        if getattr(self, "_b", None) is not None:
            _lib.lib().ffcm_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ptrs(self, xs, conv):
        return (ctypes.c_void_p * self.K)(*[conv(x) for x in xs])

    def mobility_(self, V, Y, F):
        """Host arrays: ``Y[k]`` 3 x N_k, ``F[k]`` and ``V[k]`` 3 x N_k x R (Fortran order
        preferred). Overwrites every ``V[k]``; returns ``V``."""
        T, R = self.T, self.n_rhs
        if not (len(V) == len(Y) == len(F) == self.K):
            raise DimensionMismatch(f"V, Y and F must each list {self.K} systems")
        Yc, Fc, Vo = [], [], []
        for k in range(self.K):
            N = self.N[k]
            if np.shape(Y[k]) != (3, N) or np.shape(F[k]) != (3, N, R) or np.shape(V[k]) != (3, N, R):
                raise DimensionMismatch(f"system {k}: Y must be 3x{N} and V, F 3x{N}x{R}")
            Yc.append(np.asfortranarray(Y[k], dtype=T))
            Fc.append(np.asfortranarray(F[k], dtype=T))
            v = V[k]
            Vo.append(v if (isinstance(v, np.ndarray) and v.dtype == T and v.flags.f_contiguous)
                      else np.empty((3, N, R), T, order="F"))
        data = lambda a: a.ctypes.data  # noqa: E731
        rc = _lib.lib().ffcm_batch_mobility(self._b, self._ptrs(Vo, data), self._ptrs(Yc, data), self._ptrs(Fc, data))
        _raise(rc, self._b, "batch mobility")
        for k in range(self.K):
            if Vo[k] is not V[k]:
                V[k][...] = Vo[k]
        return V

    def mobility_device_(self, V, Y, F, synchronize: bool = True):
        """Device buffers (``data_ptr()`` objects or integer addresses): ``Y[k]`` holds 3 N_k
        values, ``V[k]`` and ``F[k]`` 3 N_k R. Enqueued on the batch stream."""
        def ptr(x):
            return int(x.data_ptr()) if hasattr(x, "data_ptr") else int(x)

        R = self.n_rhs
        for k in range(self.K):
            for name, x, n in (("V", V[k], 3 * self.N[k] * R), ("Y", Y[k], 3 * self.N[k]),
                               ("F", F[k], 3 * self.N[k] * R)):
                if hasattr(x, "numel") and x.numel() != n:
                    raise DimensionMismatch(f"system {k}: {name} must hold {n} elements; got {x.numel()}")
        rc = _lib.lib().ffcm_batch_mobility_device(self._b, self._ptrs(V, ptr), self._ptrs(Y, ptr),
                                                   self._ptrs(F, ptr))
        _raise(rc, self._b, "batch mobility")
        if synchronize:
            self.synchronize()
        return V

    def synchronize(self):
        _raise(_lib.lib().ffcm_batch_synchronize(self._b), self._b, "synchronize")

    def stream(self) -> int:
        s = ctypes.c_void_p()
        _raise(_lib.lib().ffcm_batch_get_stream(self._b, ctypes.byref(s)), self._b, "stream")
        return s.value or 0

    def info(self) -> dict:
        lanes = (ctypes.c_int32 * 2)()
        nodes = ctypes.c_int64()
        tune = (ctypes.c_double * 8)()
        _raise(_lib.lib().ffcm_batch_get_info(self._b, lanes, ctypes.byref(nodes), tune), self._b, "info")
        return {"execution": "grouped" if lanes[0] < 0 else ("lanes" if lanes[0] > 0 else None),
                "lanes": abs(lanes[0]), "lanes_allocated": lanes[1], "graph_nodes": nodes.value,
                "tune_ms_per_system_by_lanes": [round(t, 4) for t in tune[:4]],
                "tune_ms_grouped_by_lanes": [round(t, 4) for t in tune[4:]]}
