// Compile-time Float32 FFT plans, part 3 of 4 (the lengths are spread over four
// translation units so that they compile in parallel). Each plan is the radix sequence of
// plan_axis (8s, then 4, 2, 3, 5) for one 5-smooth length whose threads per pencil, rounded
// up to a power of two, is at most 64.
#include "fft_kernels.hpp"

namespace ffcm {
namespace fftk {

bool static_kernels_part3(int n, FftKernels<float>* out) {
  switch (n) {
    case 60: *out = make_kernels<float, Radices<4, 3, 5>>(); return true;
    case 80: *out = make_kernels<float, Radices<8, 2, 5>>(); return true;
    case 100: *out = make_kernels<float, Radices<4, 5, 5>>(); return true;
    case 128: *out = make_kernels<float, Radices<8, 8, 2>>(); return true;
    case 160: *out = make_kernels<float, Radices<8, 4, 5>>(); return true;
    case 200: *out = make_kernels<float, Radices<8, 5, 5>>(); return true;
    case 243: *out = make_kernels<float, Radices<3, 3, 3, 3, 3>>(); return true;
    case 288: *out = make_kernels<float, Radices<8, 4, 3, 3>>(); return true;
    case 384: *out = make_kernels<float, Radices<8, 8, 2, 3>>(); return true;
    default: return false;
  }
}

}  // namespace fftk
}  // namespace ffcm
