// Compile-time Float32 FFT plans, part 0 of 4 (the lengths are spread over four
// translation units so that they compile in parallel). Each plan is the radix sequence of one
// 5-smooth length: for a power of two that of plan_axis (8s, then 4, 2); for other lengths
// the odd radices first (3s, then 5s), then 8, 4, 2, which makes the unpadded LDS layout
// bank-conflict-free from the first stage on. Each runs with its own threads per pencil
// (plan_tp, a power of two up to 64).
#include "fft_kernels.hpp"

namespace ffcm {
namespace fftk {

bool static_kernels_part0(int n, FftKernels<float>* out) {
  switch (n) {
    case 48: *out = make_kernels<float, Radices<3, 8, 2>>(); return true;
    case 64: *out = make_kernels<float, Radices<8, 8>>(); return true;
    case 81: *out = make_kernels<float, Radices<3, 3, 3, 3>>(); return true;
    case 108: *out = make_kernels<float, Radices<3, 3, 3, 4>>(); return true;
    case 135: *out = make_kernels<float, Radices<3, 3, 3, 5>>(); return true;
    case 162: *out = make_kernels<float, Radices<3, 3, 3, 3, 2>>(); return true;
    case 216: *out = make_kernels<float, Radices<3, 3, 3, 8>>(); return true;
    case 250: *out = make_kernels<float, Radices<5, 5, 5, 2>>(); return true;
    case 300: *out = make_kernels<float, Radices<3, 5, 5, 4>>(); return true;
    case 512: *out = make_kernels_wide<float, Radices<8, 8, 8>>(); return true;
    default: return false;
  }
}

}  // namespace fftk
}  // namespace ffcm
