// Compile-time Float32 FFT plans, part 2 of 4 (the lengths are spread over four
// translation units so that they compile in parallel). Each plan is the radix sequence of one
// 5-smooth length: for a power of two that of plan_axis (8s, then 4, 2); for other lengths
// the odd radices first (3s, then 5s), then 8, 4, 2, which makes the unpadded LDS layout
// bank-conflict-free from the first stage on. Each runs with its own threads per pencil
// (plan_tp, a power of two up to 64).
#include "fft_kernels.hpp"

namespace ffcm {
namespace fftk {

bool static_kernels_part2(int n, FftKernels<float>* out) {
  switch (n) {
    case 54: *out = make_kernels<float, Radices<3, 3, 3, 2>>(); return true;
    case 75: *out = make_kernels<float, Radices<3, 5, 5>>(); return true;
    case 96: *out = make_kernels<float, Radices<3, 8, 4>>(); return true;
    case 125: *out = make_kernels<float, Radices<5, 5, 5>>(); return true;
    case 150: *out = make_kernels<float, Radices<3, 5, 5, 2>>(); return true;
    case 192: *out = make_kernels<float, Radices<3, 8, 8>>(); return true;
    case 240: *out = make_kernels<float, Radices<3, 5, 8, 2>>(); return true;
    case 270: *out = make_kernels<float, Radices<3, 3, 3, 5, 2>>(); return true;
    case 324: *out = make_kernels<float, Radices<3, 3, 3, 3, 4>>(); return true;
    default: return false;
  }
}

}  // namespace fftk
}  // namespace ffcm
