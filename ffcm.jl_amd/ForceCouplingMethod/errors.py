"""Exception types mirroring the Julia exceptions the reference API throws."""


class ArgumentError(ValueError):
    """Julia ``ArgumentError``: invalid configuration (src/config.jl:267-312, :461-464)."""


class DimensionMismatch(ValueError):
    """Julia ``DimensionMismatch``: a 3xN argument of the wrong shape (src/mobility.jl:29-34)."""


class DeviceError(RuntimeError):
    """A HIP runtime failure inside the backend (CUDA.jl surfaces these as Julia exceptions)."""


class OutOfGPUMemoryError(MemoryError):
    """A device allocation failed while assembling the configuration."""
