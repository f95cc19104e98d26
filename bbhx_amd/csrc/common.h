// Shared device-side definitions for the gfx950 kernels: physical constants,
// a small complex type, status codes and launch helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace bbhx {

// Constants of the reference (cutils/constants.h, utils/constants.py).
constexpr double MSUN_SI = 1.98848e+30;
constexpr double YRSID_SI = 31558149.763545603;
constexpr double AU_SI = 149597870700.0;
constexpr double C_SI = 299792458.;
constexpr double MTSUN_SI = 4.925491025873693e-06;
constexpr double MRSUN_SI = 1476.6250615036158;
constexpr double PI = 3.141592653589793238462643383279502884;
constexpr double PI_2 = 1.570796326794896619231321691639751442;
constexpr double PI_4 = 0.785398163397448309615660845819875721;
constexpr double TWO_PI = 6.283185307179586476925286766559005768;
constexpr double GAMMA_E = 0.577215664901532860606512090082402431;
constexpr double SQRT2 = 1.414213562373095048801688724209698079;
constexpr double INVSQRT3 = 0.577350269189625764509148780501957455;
constexpr double INVSQRT6 = 0.408248290463863016366214012450981898;
constexpr double OMEGA0 = 1.9909865927683788e-07;

constexpr int MAX_MODES = 6;

// Status codes returned by every extern "C" entry point.
enum Status : int {
    OK = 0,
    ERR_ARGS = 1,      // invalid sizes / null pointers / unsupported option
    ERR_LAUNCH = 2,    // kernel launch or runtime error
    ERR_ALLOC = 3,     // workspace allocation failed
    ERR_MODE = 4,      // unsupported (l, m)
};

struct cd {
    double re, im;
};

__host__ __device__ inline cd mk(double r, double i) { return cd{r, i}; }
__host__ __device__ inline cd operator+(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cd operator-(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cd operator*(cd a, cd b) { return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__host__ __device__ inline cd operator*(double s, cd a) { return cd{s * a.re, s * a.im}; }
__host__ __device__ inline cd operator*(cd a, double s) { return cd{a.re * s, a.im * s}; }
__host__ __device__ inline cd conj(cd a) { return cd{a.re, -a.im}; }
__host__ __device__ inline cd& operator+=(cd& a, cd b) { a.re += b.re; a.im += b.im; return a; }

__device__ inline cd cexpi(double phase)
{
    double s, c;
    sincos(phase, &s, &c);
    return cd{c, s};
}

// i * a
__host__ __device__ inline cd mul_i(cd a) { return cd{-a.im, a.re}; }

inline int ceil_div(long a, long b) { return int((a + b - 1) / b); }

}  // namespace bbhx
