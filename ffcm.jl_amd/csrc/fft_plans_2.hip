// Compile-time Float32 FFT plans, part 2 of 4 (the lengths are spread over four
// translation units so that they compile in parallel). Each plan is the radix sequence of
// plan_axis (8s, then 4, 2, 3, 5) for one 5-smooth length whose threads per pencil, rounded
// up to a power of two, is at most 64.
#include "fft_kernels.hpp"

namespace ffcm {
namespace fftk {

bool static_kernels_part2(int n, FftKernels<float>* out) {
  switch (n) {
    case 54: *out = make_kernels<float, Radices<2, 3, 3, 3>>(); return true;
    case 75: *out = make_kernels<float, Radices<3, 5, 5>>(); return true;
    case 96: *out = make_kernels<float, Radices<8, 4, 3>>(); return true;
    case 125: *out = make_kernels<float, Radices<5, 5, 5>>(); return true;
    case 150: *out = make_kernels<float, Radices<2, 3, 5, 5>>(); return true;
    case 192: *out = make_kernels<float, Radices<8, 8, 3>>(); return true;
    case 240: *out = make_kernels<float, Radices<8, 2, 3, 5>>(); return true;
    case 270: *out = make_kernels<float, Radices<2, 3, 3, 3, 5>>(); return true;
    case 324: *out = make_kernels<float, Radices<4, 3, 3, 3, 3>>(); return true;
    default: return false;
  }
}

}  // namespace fftk
}  // namespace ffcm
