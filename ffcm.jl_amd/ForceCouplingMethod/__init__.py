"""Host-side mirror of ForceCouplingMethod.jl's public API over the MI355X HIP backend.

The reference exports exactly ``FFCMConfig``, ``mobility!`` and ``FFCMMobility``
(src/ForceCouplingMethod.jl:20). Here ``mobility!`` is spelled ``mobility_`` and the 5-argument
``mul!`` is ``mul_``; the step wrappers carry the same trailing underscore. Every compute call
runs the HIP kernels of libffcm_hip.so; a missing library raises ``ArgumentError``.
"""
from .batch import FFCMBatch
from .config import FFCMConfig
from .errors import ArgumentError, DeviceError, DimensionMismatch, OutOfGPUMemoryError
from .mobility import (
    FFCMMobility,
    apply_forces_,
    apply_forces_device_,
    set_positions_,
    set_positions_device_,
    assign_cells_,
    correct_velocities_,
    interpolate_velocities_,
    mobility_,
    mobility_device_,
    mobility_multi_,
    mobility_multi_device_,
    mobility_vjp_,
    mobility_vjp_device_,
    mul_,
    rrule_mobility,
    sort_particles_by_cell_,
    spread_forces_,
    stokes_backward_,
    stokes_forward_project_,
    stokes_solve_,
    torch_mobility,
    wrap_positions_,
)

__all__ = ["FFCMConfig", "mobility_", "FFCMMobility"]
