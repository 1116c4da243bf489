// Compile-time Float32 FFT plans, part 3 of 4 (the lengths are spread over four
// translation units so that they compile in parallel). Each plan is the radix sequence of one
// 5-smooth length: for a power of two that of plan_axis (8s, then 4, 2); for other lengths
// the odd radices first (3s, then 5s), then 8, 4, 2, which makes the unpadded LDS layout
// bank-conflict-free from the first stage on. Each runs with its own threads per pencil
// (plan_tp, a power of two up to 64).
#include "fft_kernels.hpp"

namespace ffcm {
namespace fftk {

bool static_kernels_part3(int n, FftKernels<float>* out) {
  switch (n) {
    case 60: *out = make_kernels<float, Radices<3, 5, 4>>(); return true;
    case 80: *out = make_kernels<float, Radices<5, 8, 2>>(); return true;
    case 100: *out = make_kernels<float, Radices<5, 5, 4>>(); return true;
    case 128: *out = make_kernels<float, Radices<8, 8, 2>>(); return true;
    case 160: *out = make_kernels<float, Radices<5, 8, 4>>(); return true;
    case 200: *out = make_kernels<float, Radices<5, 5, 8>>(); return true;
    case 243: *out = make_kernels<float, Radices<3, 3, 3, 3, 3>>(); return true;
    case 288: *out = make_kernels<float, Radices<3, 3, 8, 4>>(); return true;
    case 384: *out = make_kernels<float, Radices<3, 8, 8, 2>>(); return true;
    default: return false;
  }
}

}  // namespace fftk
}  // namespace ffcm
