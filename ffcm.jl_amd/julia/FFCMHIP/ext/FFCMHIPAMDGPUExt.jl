# Zero-copy mobility! for device arrays (AMDGPU.jl ROCArrays): the device branch of the staging
# seam, which the reference's CUDA extension serves with a device -> device copyto!
# (ext/ForceCouplingMethodCUDAExt/mobility.jl:1-16). Here the caller's device arrays go to
# ffcm_mobility_device directly, with no staging copy.
module FFCMHIPAMDGPUExt

using AMDGPU: ROCMatrix
import AMDGPU
using FFCMHIP: FFCMHIP, HIPConfig, mobility_device!, synchronize
import ForceCouplingMethod: mobility!

function mobility!(V::ROCMatrix{T}, config::HIPConfig{T}, Y::ROCMatrix{T}, F::ROCMatrix{T}) where {T}
    N = size(config.particles.Y_sorted, 2)
    (size(V) == (3, N) && size(Y) == (3, N) && size(F) == (3, N)) || throw(DimensionMismatch(
        "V, Y, and F must each be 3xN for the config's N = $(N)"))
    AMDGPU.synchronize()  # the caller's writes to Y and F, on AMDGPU.jl's stream, are done
    mobility_device!(pointer(V), config, pointer(Y), pointer(F))
    synchronize(config)   # V is complete before AMDGPU.jl's stream reads it
    return V
end

end # module FFCMHIPAMDGPUExt
