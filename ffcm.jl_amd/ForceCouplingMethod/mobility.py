"""``mobility!``, the step wrappers and the matrix-free ``FFCMMobility`` operator.

Host-side mirror of src/mobility.jl:23-162 and of the L2 step wrappers of src/cell_list.jl,
src/spread_forces.jl, src/stokes_solve.jl, src/interpolate.jl and src/correct_velocities.jl.
Every call goes to the HIP backend; there is no other compute path.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .config import FFCMConfig
from .errors import ArgumentError, DimensionMismatch


def _as_3xN(x, T, N, name):
    a = np.asarray(x)
    if a.shape != (3, N):
        raise DimensionMismatch(
            f"V, Y, and F must each be 3xN for the config's N = {N}; got size({name}) = {a.shape}"
        )
    return np.asfortranarray(a, dtype=T)


def mobility_(V, config: FFCMConfig, Y, F):
    """``mobility!(V, config, Y, F)`` (src/mobility.jl:23-45): overwrite the 3xN ``V`` with the
    velocities induced by forces ``F`` at positions ``Y`` (all in the caller's original particle
    order). ``Y`` and ``F`` are read only. Returns ``V``."""
    T, N = config.T, config.N
    for name, x in (("V", V), ("Y", Y), ("F", F)):
        if np.shape(x) != (3, N):
            raise DimensionMismatch(
                f"V, Y, and F must each be 3xN for the config's N = {N}; got size(V) = {np.shape(V)}, "
                f"size(Y) = {np.shape(Y)}, size(F) = {np.shape(F)}"
            )
    Yc = _as_3xN(Y, T, N, "Y")
    Fc = _as_3xN(F, T, N, "F")
    out = V if (isinstance(V, np.ndarray) and V.dtype == T and V.flags.f_contiguous) else np.empty((3, N), T, order="F")
    config._positions_owner = None
    config._check(
        _lib.lib().ffcm_mobility(config.ctx, out.ctypes.data, Yc.ctypes.data, Fc.ctypes.data, N), "mobility!"
    )
    if out is not V:
        V[...] = out
    return V


def _dptr(x):
    return int(x.data_ptr()) if hasattr(x, "data_ptr") else int(x)


def set_positions_(config: FFCMConfig, Y, owner=None):
    """Cache steps 1-2 of ``mobility!`` (wrap, hash, sort, binning) for the 3xN host positions
    ``Y`` in the config's context; ``apply_forces_`` then runs the rest for any forces. The
    reference rebuilds the cell list on every ``mobility!`` (src/mobility.jl:35-43). ``owner``
    records who cached them (an ``FFCMMobility``)."""
    T, N = config.T, config.N
    Yc = _as_3xN(Y, T, N, "Y")
    config._check(_lib.lib().ffcm_set_positions(config.ctx, Yc.ctypes.data, N), "set_positions!")
    config._positions_owner = owner
    return config


def set_positions_device_(config: FFCMConfig, Y, owner=None, synchronize: bool = True):
    """``set_positions_`` for a device buffer of 3N elements of T (enqueued on the config's stream)."""
    if hasattr(Y, "numel") and Y.numel() != 3 * config.N:
        raise DimensionMismatch(f"Y must hold 3xN = {3 * config.N} elements; got {Y.numel()}")
    config._check(_lib.lib().ffcm_set_positions_device(config.ctx, _dptr(Y), config.N), "set_positions!")
    config._positions_owner = owner
    if synchronize:
        config.synchronize()
    return config


def apply_forces_(V, config: FFCMConfig, F):
    """``mobility!(V, config, Y, F)`` for the ``Y`` of the last ``set_positions_`` (host 3xN)."""
    T, N = config.T, config.N
    Fc = _as_3xN(F, T, N, "F")
    if np.shape(V) != (3, N):
        raise DimensionMismatch(f"V must be 3xN for the config's N = {N}; got size(V) = {np.shape(V)}")
    out = V if (isinstance(V, np.ndarray) and V.dtype == T and V.flags.f_contiguous) else np.empty((3, N), T, order="F")
    config._check(_lib.lib().ffcm_apply_forces(config.ctx, out.ctypes.data, Fc.ctypes.data, N), "mul!")
    if out is not V:
        V[...] = out
    return V


def apply_forces_device_(V, config: FFCMConfig, F, synchronize: bool = True):
    """``apply_forces_`` on device buffers of 3N elements of T (enqueued on the config's stream)."""
    for name, x in (("V", V), ("F", F)):
        if hasattr(x, "numel") and x.numel() != 3 * config.N:
            raise DimensionMismatch(f"{name} must hold 3xN = {3 * config.N} elements; got {x.numel()}")
    config._check(_lib.lib().ffcm_apply_forces_device(config.ctx, _dptr(V), _dptr(F), config.N), "mul!")
    if synchronize:
        config.synchronize()
    return V


def mobility_device_(V, config: FFCMConfig, Y, F, synchronize: bool = True):
    """Device-resident ``mobility!``: ``V``, ``Y``, ``F`` are device buffers of 3N elements of T
    (objects with ``data_ptr()``, or integer device addresses), the zero-copy branch of the
    staging seam (ext/ForceCouplingMethodCUDAExt/mobility.jl:1-8). Enqueued on the config's
    stream; synchronises only when asked."""
    def ptr(x):
        return int(x.data_ptr()) if hasattr(x, "data_ptr") else int(x)

    for name, x in (("V", V), ("Y", Y), ("F", F)):
        if hasattr(x, "numel") and x.numel() != 3 * config.N:
            raise DimensionMismatch(f"{name} must hold 3xN = {3 * config.N} elements; got {x.numel()}")
    config._positions_owner = None
    config._check(_lib.lib().ffcm_mobility_device(config.ctx, ptr(V), ptr(Y), ptr(F), config.N), "mobility!")
    if synchronize:
        config.synchronize()
    return V


# ------------------------------------------------------------------ several right-hand sides
def mobility_multi_(V, config: FFCMConfig, Y, F):
    """R right-hand sides sharing ``Y``: ``V[:, :, r] = M(Y) F[:, :, r]`` for 3 x N x R arrays
    (R <= the config's ``max_rhs``). Each right-hand side equals a separate ``mobility!``; the
    cell list is built once and the R force grids are transformed together. Returns ``V``."""
    T, N = config.T, config.N
    F_ = np.asarray(F)
    if F_.ndim != 3 or F_.shape[:2] != (3, N) or np.shape(V) != F_.shape or np.shape(Y) != (3, N):
        raise DimensionMismatch(f"V and F must be 3 x N x R and Y 3 x N for the config's N = {N}; got "
                                f"size(V) = {np.shape(V)}, size(Y) = {np.shape(Y)}, size(F) = {F_.shape}")
    R = F_.shape[2]
    Yc = _as_3xN(Y, T, N, "Y")
    Fc = np.asfortranarray(F_, dtype=T)
    out = V if (isinstance(V, np.ndarray) and V.dtype == T and V.flags.f_contiguous) else np.empty((3, N, R), T, order="F")
    config._positions_owner = None
    config._check(_lib.lib().ffcm_mobility_multi(config.ctx, R, out.ctypes.data, Yc.ctypes.data, Fc.ctypes.data, N),
                  "mobility_multi")
    if out is not V:
        V[...] = out
    return V


def mobility_multi_device_(V, config: FFCMConfig, Y, F, R: int, synchronize: bool = True):
    """Device-resident ``mobility_multi_``: ``V`` and ``F`` hold R consecutive 3N blocks, ``Y``
    one (``data_ptr()`` objects or integer addresses). Enqueued on the config's stream."""
    def ptr(x):
        return int(x.data_ptr()) if hasattr(x, "data_ptr") else int(x)

    for name, x, n in (("V", V, 3 * config.N * R), ("Y", Y, 3 * config.N), ("F", F, 3 * config.N * R)):
        if hasattr(x, "numel") and x.numel() != n:
            raise DimensionMismatch(f"{name} must hold {n} elements; got {x.numel()}")
    config._positions_owner = None
    config._check(_lib.lib().ffcm_mobility_multi_device(config.ctx, int(R), ptr(V), ptr(Y), ptr(F), config.N),
                  "mobility_multi")
    if synchronize:
        config.synchronize()
    return V


# ------------------------------------------------------------------ adjoint (rrule)
def mobility_vjp_(dF, dY, config: FFCMConfig, Y, F, G):
    """Pullback of ``mobility!`` for L = <G, M(Y) F>: overwrites the 3xN ``dF`` with
    dL/dF = M(Y) G and ``dY`` with dL/dY (gradient-of-Gaussian stencil weights plus the
    r-derivative of the pair correction). The cell list and weights are recomputed from ``Y``.
    ``dY`` is the derivative of the smooth pieces of V(Y): it does not include the jumps where
    a stencil anchor round(Y/h) switches or a pair crosses r = R_c. Returns ``(dF, dY)``."""
    T, N = config.T, config.N
    for name, x in (("dF", dF), ("dY", dY), ("Y", Y), ("F", F), ("G", G)):
        if np.shape(x) != (3, N):
            raise DimensionMismatch(f"dF, dY, Y, F and G must each be 3xN for the config's N = {N}; "
                                    f"got size({name}) = {np.shape(x)}")
    Yc, Fc, Gc = (_as_3xN(x, T, N, n) for x, n in ((Y, "Y"), (F, "F"), (G, "G")))

    def out_of(x):
        return x if (isinstance(x, np.ndarray) and x.dtype == T and x.flags.f_contiguous) else np.empty((3, N), T, order="F")

    oF, oY = out_of(dF), out_of(dY)
    config._positions_owner = None
    config._check(
        _lib.lib().ffcm_mobility_vjp(config.ctx, oF.ctypes.data, oY.ctypes.data, Yc.ctypes.data, Fc.ctypes.data,
                                     Gc.ctypes.data, N),
        "mobility_vjp",
    )
    if oF is not dF:
        dF[...] = oF
    if oY is not dY:
        dY[...] = oY
    return dF, dY


def mobility_vjp_device_(dF, dY, config: FFCMConfig, Y, F, G, synchronize: bool = True):
    """Device-resident ``mobility_vjp_``: all five arguments are device buffers of 3N elements
    of T (``data_ptr()`` objects or integer addresses). Enqueued on the config's stream."""
    def ptr(x):
        return int(x.data_ptr()) if hasattr(x, "data_ptr") else int(x)

    for name, x in (("dF", dF), ("dY", dY), ("Y", Y), ("F", F), ("G", G)):
        if hasattr(x, "numel") and x.numel() != 3 * config.N:
            raise DimensionMismatch(f"{name} must hold 3xN = {3 * config.N} elements; got {x.numel()}")
    config._positions_owner = None
    config._check(
        _lib.lib().ffcm_mobility_vjp_device(config.ctx, ptr(dF), ptr(dY), ptr(Y), ptr(F), ptr(G), config.N),
        "mobility_vjp",
    )
    if synchronize:
        config.synchronize()
    return dF, dY


def rrule_mobility(config: FFCMConfig, Y, F):
    """The ChainRulesCore ``rrule`` of ``mobility!`` as a pair (V, pullback): V = M(Y) F and
    ``pullback(G) -> (dY, dF)`` with the gradients of <G, V>. Nothing from the forward call is
    kept except Y and F; the pullback recomputes the cell list and weights."""
    T, N = config.T, config.N
    Yc = np.array(_as_3xN(Y, T, N, "Y"), order="F")
    Fc = np.array(_as_3xN(F, T, N, "F"), order="F")
    V = mobility_(np.empty((3, N), T, order="F"), config, Yc, Fc)

    def pullback(G):
        dF = np.empty((3, N), T, order="F")
        dY = np.empty((3, N), T, order="F")
        mobility_vjp_(dF, dY, config, Yc, Fc, G)
        return dY, dF

    return V, pullback


def torch_mobility(config: FFCMConfig, Y, F):
    """``V = M(Y) F`` as a differentiable torch operation on device tensors (3N elements each,
    column = particle). Backward calls the pullback kernels."""
    import torch

    class _Mobility(torch.autograd.Function):
        @staticmethod
        def forward(ctx, Yt, Ft):
            Yt = Yt.contiguous()
            Ft = Ft.contiguous()
            V = torch.empty_like(Ft)
            mobility_device_(V, config, Yt, Ft)
            ctx.save_for_backward(Yt, Ft)
            return V

        @staticmethod
        def backward(ctx, Gt):
            Yt, Ft = ctx.saved_tensors
            Gt = Gt.contiguous()
            dF = torch.empty_like(Ft)
            dY = torch.empty_like(Yt)
            mobility_vjp_device_(dF, dY, config, Yt, Ft, Gt)
            return dY, dF

    return _Mobility.apply(Y, F)


# ------------------------------------------------------------------ step wrappers (L2)
def _step(fn_name, what):
    def step(config: FFCMConfig):
        config._positions_owner = None
        config._check(getattr(_lib.lib(), fn_name)(config.ctx), what)
        return config

    step.__name__ = what
    return step


wrap_positions_ = _step("ffcm_wrap_positions", "wrap_positions!")  # src/cell_list.jl:22-31
assign_cells_ = _step("ffcm_assign_cells", "assign_cells!")  # src/cell_list.jl:193-198
sort_particles_by_cell_ = _step("ffcm_sort_particles_by_cell", "sort_particles_by_cell!")  # :119-133
spread_forces_ = _step("ffcm_spread_forces", "spread_forces!")  # src/spread_forces.jl:17-33
stokes_solve_ = _step("ffcm_stokes_solve", "stokes_solve!")  # src/stokes_solve.jl:21-42
stokes_forward_project_ = _step("ffcm_stokes_forward_project", "stokes_forward_project!")
stokes_backward_ = _step("ffcm_stokes_backward", "stokes_backward!")
interpolate_velocities_ = _step("ffcm_interpolate_velocities", "interpolate_velocities!")  # src/interpolate.jl:23-42
correct_velocities_ = _step("ffcm_correct_velocities", "correct_velocities!")  # src/correct_velocities.jl:154-173


class FFCMMobility:
    """Matrix-free 3N x 3N mobility operator at fixed positions (src/mobility.jl:88-162).

    The flat vector index of particle n, axis i is 3n + i (0-based), i.e. the column-major
    flattening of a 3xN matrix. ``mul_(v, M, f)`` and ``mul_(v, M, f, alpha, beta)`` are the
    3- and 5-argument ``mul!``; ``M @ F`` applies it to a 3xN matrix (``*``)."""

    def __init__(self, config: FFCMConfig, Y):
        Y = np.asarray(Y)
        if Y.ndim != 2 or Y.shape[0] != 3:
            raise ArgumentError(f"Y must be 3xN; got size {Y.shape}")
        N = Y.shape[1]
        if N != config.N:
            raise ArgumentError(
                f"Y has {N} particles but config was built for {config.N}; construct a config with matching N"
            )
        self.config = config
        self.Y = np.asfortranarray(Y, dtype=config.T)
        self.F_scratch = np.empty((3, N), config.T, order="F")
        self.V_scratch = np.empty((3, N), config.T, order="F")
        self._Y_cached = None  # the positions whose cell list the config holds for this operator

    def _apply(self, V, F):
        """mobility!(V, config, Y, F) for this operator's Y: the cell list of Y stays cached in
        the config between applies (steps 1-2 run again when Y changed in place or another
        caller of the same config replaced the cell list)."""
        if (getattr(self.config, "_positions_owner", None) is not self or self._Y_cached is None
                or not np.array_equal(self._Y_cached, self.Y)):
            set_positions_(self.config, self.Y, owner=self)
            self._Y_cached = self.Y.copy()
        return apply_forces_(V, self.config, F)

    @property
    def shape(self):
        n = 3 * self.Y.shape[1]
        return (n, n)

    def size(self, d=None):
        n = 3 * self.Y.shape[1]
        if d is None:
            return (n, n)
        return n if d <= 2 else 1

    @property
    def dtype(self):
        return np.dtype(self.config.T)

    def eltype(self):
        return self.config.T

    def mul_(self, v, f, alpha=None, beta=None):
        self.F_scratch.reshape(-1, order="F")[:] = np.asarray(f).reshape(-1)
        self._apply(self.V_scratch, self.F_scratch)
        res = self.V_scratch.reshape(-1, order="F")
        if alpha is None and beta is None:
            v[...] = res
            return v
        T = self.config.T
        a = T(1 if alpha is None else alpha)
        b = T(0 if beta is None else beta)
        if b == 0:
            v[...] = a * res  # beta = 0 overwrites v, ignoring its contents
        else:
            v[...] = a * res + b * np.asarray(v)
        return v

    def __matmul__(self, F):
        N = self.Y.shape[1]
        if np.shape(F) != (3, N):
            raise DimensionMismatch(f"V, Y, and F must each be 3xN for the config's N = {N}; got size(F) = {np.shape(F)}")
        return self._apply(np.empty((3, N), self.config.T, order="F"), F)

    __mul__ = __matmul__

    def issymmetric(self):
        return True

    def isposdef(self):
        return True


def mul_(v, M: FFCMMobility, f, alpha=None, beta=None):
    """``mul!(v, M, f[, α, β])`` (src/mobility.jl:115-147)."""
    return M.mul_(v, f, alpha, beta)
