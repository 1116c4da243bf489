"""ctypes bindings of libffcm_hip.so (include/ffcm_hip.h).

The library is the only compute path of this package: nothing here falls back to a CPU
implementation. A missing or unloadable library raises ``ArgumentError``, the error the
reference raises when ``gpu_acceleration = true`` has no backend (src/config.jl:461-464).
"""
from __future__ import annotations

import ctypes
import os

from .errors import ArgumentError

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FFCM_HIP_LIB", os.path.join(os.path.dirname(_PKG), "libffcm_hip.so"))

FFCM_OK = 0
FFCM_ERR_ARGUMENT = 1
FFCM_ERR_DIMENSION = 2
FFCM_ERR_DEVICE = 3
FFCM_ERR_MEMORY = 4
FFCM_F32 = 0
FFCM_F64 = 1

BUFFER_IDS = {
    "Y_input": 0,
    "F_input": 1,
    "V_output": 2,
    "Y_wrapped": 3,
    "Y_sorted": 4,
    "F_sorted": 5,
    "cell_hash": 6,
    "original_index": 7,
    "cell_start": 8,
    "cell_end": 9,
    "force_density": 10,
    "fluid_velocity": 11,
    "fluid_hat": 12,
    "neighbor_map": 13,
    "counting_sort_scratch": 14,
}

STAGES = ("cell_list", "spread", "stokes_fft", "interpolate", "correct")

# Every symbol include/ffcm_hip.h declares.
EXPORTS = (
    "ffcm_create", "ffcm_destroy", "ffcm_last_error", "ffcm_mobility", "ffcm_mobility_device",
    "ffcm_wrap_positions", "ffcm_assign_cells", "ffcm_sort_particles_by_cell", "ffcm_spread_forces",
    "ffcm_stokes_solve", "ffcm_interpolate_velocities", "ffcm_correct_velocities",
    "ffcm_stokes_forward_project", "ffcm_stokes_backward", "ffcm_buffer_bytes", "ffcm_set_buffer",
    "ffcm_get_buffer", "ffcm_get_stream", "ffcm_synchronize", "ffcm_get_info", "ffcm_set_timing", "ffcm_set_graph",
    "ffcm_get_timing", "ffcm_mobility_vjp", "ffcm_mobility_vjp_device", "ffcm_create_multi",
    "ffcm_mobility_multi", "ffcm_mobility_multi_device", "ffcm_batch_create", "ffcm_batch_destroy",
    "ffcm_batch_last_error", "ffcm_batch_mobility", "ffcm_batch_mobility_device", "ffcm_batch_get_stream",
    "ffcm_batch_synchronize", "ffcm_batch_get_info", "ffcm_set_positions", "ffcm_set_positions_device",
    "ffcm_apply_forces", "ffcm_apply_forces_device", "ffcm_assemble", "ffcm_set_scalars",
    "ffcm_wrap_positions_kernel", "ffcm_assign_cells_kernel", "ffcm_build_cell_list_kernel",
    "ffcm_gather_particles_kernel", "ffcm_spread_forces_kernel", "ffcm_forward_transform",
    "ffcm_apply_inverse_stokes_kernel", "ffcm_inverse_transform", "ffcm_interpolate_velocities_kernel",
    "ffcm_correct_velocities_kernel", "ffcm_mobility_device_noalloc", "ffcm_noalloc_last_error",
    "ffcm_noalloc_scratch_bytes",
)


class Params(ctypes.Structure):
    """``ffcm_params`` (include/ffcm_hip.h)."""

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
    """Loads the library once; raises ArgumentError when it is not built or cannot load."""
    global _lib
    if _lib is not None:
        return _lib
    # One HIP runtime per process: when PyTorch (which bundles its own libamdhip64) is
    # installed, load it first so this library binds to the same runtime instead of loading a
    # second copy that would leave torch unable to see the GPU.
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    if not os.path.exists(LIB_PATH):
        raise ArgumentError(
            f"gpu_acceleration = true needs the HIP backend library, which is not built: {LIB_PATH}"
        )
    try:
        L = ctypes.CDLL(LIB_PATH)
    except OSError as e:  # pragma: no cover - depends on the ROCm runtime install
        raise ArgumentError(f"cannot load the HIP backend library {LIB_PATH}: {e}") from e
    vp, i32, i64, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_size_t
    L.ffcm_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Params)]
    L.ffcm_destroy.argtypes = [vp]
    L.ffcm_last_error.argtypes = [vp]
    L.ffcm_last_error.restype = ctypes.c_char_p
    for name in ("ffcm_mobility", "ffcm_mobility_device"):
        getattr(L, name).argtypes = [vp, vp, vp, vp, i64]
    dp, ip, dbl = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32), ctypes.c_double
    L.ffcm_assemble.argtypes = [ctypes.POINTER(vp), i32, i64, ip, i32, i64, vp, vp, vp, ip]
    L.ffcm_set_scalars.argtypes = [vp, ctypes.POINTER(Params)]
    L.ffcm_wrap_positions_kernel.argtypes = [vp, i32, i32, dp]
    L.ffcm_assign_cells_kernel.argtypes = [vp, i32, i32, dp, ip]
    L.ffcm_build_cell_list_kernel.argtypes = [vp]
    L.ffcm_gather_particles_kernel.argtypes = [vp, i32, i32]
    L.ffcm_spread_forces_kernel.argtypes = [vp, dbl, dbl, dbl, dbl, ip, i32]
    L.ffcm_forward_transform.argtypes = [vp, i32]
    L.ffcm_apply_inverse_stokes_kernel.argtypes = [vp, dbl, dbl]
    L.ffcm_inverse_transform.argtypes = [vp, i32]
    L.ffcm_interpolate_velocities_kernel.argtypes = [vp, i32, dbl, dbl, dbl, dbl, ip, i32]
    L.ffcm_correct_velocities_kernel.argtypes = [vp, i32, ip, dp, dp, dbl, dbl, dbl, dbl, dbl]
    for name in ("ffcm_set_positions", "ffcm_set_positions_device"):
        getattr(L, name).argtypes = [vp, vp, i64]
    for name in ("ffcm_apply_forces", "ffcm_apply_forces_device"):
        getattr(L, name).argtypes = [vp, vp, vp, i64]
    L.ffcm_create_multi.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Params), i32]
    for name in ("ffcm_mobility_multi", "ffcm_mobility_multi_device"):
        getattr(L, name).argtypes = [vp, i32, vp, vp, vp, i64]
    L.ffcm_batch_create.argtypes = [ctypes.POINTER(vp), i32, ctypes.POINTER(Params), i32]
    L.ffcm_batch_destroy.argtypes = [vp]
    L.ffcm_batch_last_error.argtypes = [vp]
    L.ffcm_batch_last_error.restype = ctypes.c_char_p
    for name in ("ffcm_batch_mobility", "ffcm_batch_mobility_device"):
        getattr(L, name).argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp)]
    L.ffcm_batch_get_stream.argtypes = [vp, ctypes.POINTER(vp)]
    L.ffcm_batch_synchronize.argtypes = [vp]
    L.ffcm_batch_get_info.argtypes = [vp, vp, ctypes.POINTER(i64), vp]
    for name in ("ffcm_mobility_vjp", "ffcm_mobility_vjp_device"):
        getattr(L, name).argtypes = [vp, vp, vp, vp, vp, vp, i64]
    for name in ("ffcm_wrap_positions", "ffcm_assign_cells", "ffcm_sort_particles_by_cell",
                 "ffcm_spread_forces", "ffcm_stokes_solve", "ffcm_interpolate_velocities",
                 "ffcm_correct_velocities", "ffcm_stokes_forward_project", "ffcm_stokes_backward",
                 "ffcm_synchronize"):
        getattr(L, name).argtypes = [vp]
    L.ffcm_buffer_bytes.argtypes = [vp, i32, ctypes.POINTER(sz)]
    L.ffcm_set_buffer.argtypes = [vp, i32, vp, sz]
    L.ffcm_get_buffer.argtypes = [vp, i32, vp, sz]
    L.ffcm_get_stream.argtypes = [vp, ctypes.POINTER(vp)]
    L.ffcm_get_info.argtypes = [vp, vp, vp, vp, vp]
    L.ffcm_set_timing.argtypes = [vp, i32]
    L.ffcm_set_graph.argtypes = [vp, i32]
    L.ffcm_get_timing.argtypes = [vp, vp, i32, ctypes.POINTER(i64)]
    L.ffcm_mobility_device_noalloc.argtypes = [ctypes.POINTER(Params), vp, vp, vp, i64, vp]
    L.ffcm_noalloc_last_error.argtypes = []
    L.ffcm_noalloc_last_error.restype = ctypes.c_char_p
    L.ffcm_noalloc_scratch_bytes.argtypes = []
    L.ffcm_noalloc_scratch_bytes.restype = ctypes.c_size_t
    for name in EXPORTS:
        if name not in ("ffcm_last_error", "ffcm_batch_last_error", "ffcm_noalloc_last_error",
                        "ffcm_noalloc_scratch_bytes"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def last_error(ctx) -> str:
    msg = lib().ffcm_last_error(ctx)
    return msg.decode() if msg else ""
