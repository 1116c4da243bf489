// Compile-time Float32 FFT plans, part 1 of 4 (the lengths are spread over four
// translation units so that they compile in parallel). Each plan is the radix sequence of
// plan_axis (8s, then 4, 2, 3, 5) for one 5-smooth length whose threads per pencil, rounded
// up to a power of two, is at most 64.
#include "fft_kernels.hpp"

namespace ffcm {
namespace fftk {

bool static_kernels_part1(int n, FftKernels<float>* out) {
  switch (n) {
    case 50: *out = make_kernels<float, Radices<2, 5, 5>>(); return true;
    case 72: *out = make_kernels<float, Radices<8, 3, 3>>(); return true;
    case 90: *out = make_kernels<float, Radices<2, 3, 3, 5>>(); return true;
    case 120: *out = make_kernels<float, Radices<8, 3, 5>>(); return true;
    case 144: *out = make_kernels<float, Radices<8, 2, 3, 3>>(); return true;
    case 180: *out = make_kernels<float, Radices<4, 3, 3, 5>>(); return true;
    case 225: *out = make_kernels<float, Radices<3, 3, 5, 5>>(); return true;
    case 256: *out = make_kernels<float, Radices<8, 8, 4>>(); return true;
    case 320: *out = make_kernels<float, Radices<8, 8, 5>>(); return true;
    default: return false;
  }
}

}  // namespace fftk
}  // namespace ffcm
