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
constexpr double LOG2_10 = 3.321928094887362347870319429489390176;
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
    ERR_CAPACITY = 5,  // batch larger than a fixed-capacity (allocation-free) entry point holds
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

// theta mod 2 pi in two Cody-Waite steps (2 pi = HI + LO to ~107 bits); exact
// to the rounding of theta itself for |theta| up to ~1e13.
__device__ inline double reduce_2pi(double theta)
{
    const double TWO_PI_HI = 6.283185307179586;
    const double TWO_PI_LO = 2.4492935982947064e-16;
    const double INV_TWO_PI = 0.15915494309189535;
    double k = rint(theta * INV_TWO_PI);
    double r = fma(-k, TWO_PI_HI, theta);
    return fma(-k, TWO_PI_LO, r);
}

// 1/x for finite normal x != 0: the hardware reciprocal estimate refined by
// two Newton steps e = 1 - x r, r += r e (within ~1 ulp of the rounded quotient).
__device__ inline double rcp_pos(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// x^(1/3) for finite x > 0 in the single-precision range (all uses here are
// frequencies and mass ratios): a single-precision estimate 2^(log2(x)/3)
// (relative error ~1e-6) and one Halley step y (y^3 + 2x) / (2 y^3 + x), whose
// cubic convergence leaves only FP64 rounding (within ~2 ulp of cbrt).
__device__ inline double cbrt_pos(double x)
{
    const float yf = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf((float)x) * (1.0f / 3.0f));
    const double y = (double)yf;
    const double y3 = y * y * y;
    return y * (y3 + 2.0 * x) * rcp_pos(2.0 * y3 + x);
}

// x^(1/6) for finite x > 0 in the single-precision range: a single-precision
// estimate 2^(log2(x)/6) (relative error ~2e-7) and one Halley step for
// y^6 = x, y (5 y^6 + 7 x) / (7 y^6 + 5 x), whose cubic convergence leaves
// only FP64 rounding (within a few ulp of pow(x, 1/6)); one root in place of
// a square root followed by a cube root.
__device__ inline double root6_pos(double x)
{
    const float yf = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf((float)x) * (1.0f / 6.0f));
    const double y = (double)yf;
    const double y2 = y * y;
    const double y6 = y2 * y2 * y2;
    return y * fma(5.0, y6, 7.0 * x) * rcp_pos(fma(7.0, y6, 5.0 * x));
}

// Natural logarithm of a finite normal x > 0. x = 2^e m with m in
// [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.1716,
// from the atanh Taylor series 2 (s + s^3/3 + ... + s^21/21) (truncation below
// 3e-18 relative: s^22 / 23 < 2e-19); log x = e ln2 + log m with ln2 in two
// parts. m - 1 is exact (Sterbenz), s is corrected once after the reciprocal,
// and the result is within ~1 ulp of the correctly rounded log.
__device__ inline double log_pos(double x)
{
    double m = __builtin_amdgcn_frexp_mant(x);  // [0.5, 1)
    int e = __builtin_amdgcn_frexp_exp(x);
    if (m < 0.70710678118654752) {
        m = m + m;
        e -= 1;
    }
    const double num = m - 1.0, den = m + 1.0;
    const double r = rcp_pos(den);
    double s = num * r;
    s = fma(fma(-s, den, num), r, s);
    const double t = s * s;
    double q = 1.0 / 21.0;
    q = fma(q, t, 1.0 / 19.0);
    q = fma(q, t, 1.0 / 17.0);
    q = fma(q, t, 1.0 / 15.0);
    q = fma(q, t, 1.0 / 13.0);
    q = fma(q, t, 1.0 / 11.0);
    q = fma(q, t, 1.0 / 9.0);
    q = fma(q, t, 1.0 / 7.0);
    q = fma(q, t, 1.0 / 5.0);
    q = fma(q, t, 1.0 / 3.0);
    const double s2 = s + s;
    const double lm = fma(s2 * t, q, s2);
    const double ed = (double)e;
    const double LN2_HI = 0.69314718055994529, LN2_LO = 2.3190468138462996e-17;
    return fma(ed, LN2_HI, fma(ed, LN2_LO, lm));
}

// e^x for finite x: x = k ln2 + r with ln2 in the two parts of fdlibm's
// Cody-Waite reduction (|r| <= ln2 / 2), e^r by its Taylor series to r^13/13!
// (truncation < 5e-18 relative), then 2^k by the exponent (0 below the
// double range). Within ~1 ulp.
__device__ inline double exp_fast(double x)
{
    const double k = rint(x * 1.4426950408889634);
    double r = fma(-k, 6.93147180369123816490e-01, x);
    r = fma(-k, 1.90821492927058770002e-10, r);
    double p = 1.6059043836821613e-10;
    p = fma(p, r, 2.08767569878681e-09);
    p = fma(p, r, 2.505210838544172e-08);
    p = fma(p, r, 2.755731922398589e-07);
    p = fma(p, r, 2.7557319223985893e-06);
    p = fma(p, r, 2.48015873015873e-05);
    p = fma(p, r, 0.0001984126984126984);
    p = fma(p, r, 0.001388888888888889);
    p = fma(p, r, 0.008333333333333333);
    p = fma(p, r, 0.041666666666666664);
    p = fma(p, r, 0.16666666666666666);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return __builtin_amdgcn_ldexp(p, (int)k);
}

// arctan of a finite x: |x| > 1 by atan|x| = pi/2 - atan(1/|x|); y in [0, 1]
// then by the addition formula atan y = atan c_j + atan((y - c_j) / (1 + y c_j))
// with c_j the double nearest tan(j pi/12), j = 0..3 the nearest node
// (|argument| <= tan(pi/24) = 0.1317), atan c_j in two parts (j pi/12 plus the
// first-order shift of the rounded c_j) and the last term by its Taylor series
// to r^17 (truncation < 1e-18 relative); pi/2 in two parts. Within ~2 ulp.
__device__ inline double atan_fast(double x)
{
    const double a = fabs(x);
    const bool inv = a > 1.0;
    const double y = inv ? rcp_pos(a) : a;
    const int j = (y > 0.13165249758739583) + (y > 0.41421356237309503) + (y > 0.7673269879789604);
    const double c = j == 0 ? 0.0 : (j == 1 ? 0.2679491924311227 : (j == 2 ? 0.5773502691896257 : 1.0));
    const double bh = j == 0 ? 0.0 : (j == 1 ? 0.2617993877991494 : (j == 2 ? 0.5235987755982988 : 0.7853981633974483));
    const double bl = j == 0 ? 0.0 : (j == 1 ? 1.8752499114174e-17 : (j == 2 ? 3.2330503585443845e-17 : 3.061616997868383e-17));
    const double num = y - c, den = fma(y, c, 1.0);
    const double id = rcp_pos(den);
    double r = num * id;
    r = fma(fma(-r, den, num), id, r);
    const double r2 = r * r;
    double q = 1.0 / 17.0;
    q = fma(q, r2, -1.0 / 15.0);
    q = fma(q, r2, 1.0 / 13.0);
    q = fma(q, r2, -1.0 / 11.0);
    q = fma(q, r2, 1.0 / 9.0);
    q = fma(q, r2, -1.0 / 7.0);
    q = fma(q, r2, 1.0 / 5.0);
    q = fma(q, r2, -1.0 / 3.0);
    const double at = bh + (bl + fma(r * r2, q, r));
    const double res = inv ? 1.5707963267948966 - (at - 6.123233995736766e-17) : at;
    return copysign(res, x);
}

// A polynomial coefficient as a scalar-register operand at its point of use
// (rebuilt there by two scalar moves): the FMA chains below then take it as
// their addend directly, instead of copying a vector-register copy of the
// constant into the accumulator before each fmac.
__device__ inline double scoef(double c)
{
    asm volatile("" : "+s"(c));
    return c;
}

// sin and cos for |x| <= ~4 (e.g. after reduce_2pi): quadrant reduction by
// pi/2 = HI + LO, then the Taylor series of sin (to x^17) and cos (to x^18)
// on |r| <= pi/4, whose truncation errors (< 1e-19) are below FP64 rounding.
// Small and branch-light, unlike a general libm sincos with its large-argument path.
__device__ inline void sincos_small(double x, double* s, double* c)
{
    const double PIO2_HI = 1.5707963267948966;
    const double PIO2_LO = 6.123233995736766e-17;
    const double TWO_OVER_PI = 0.6366197723675814;
    const double q = rint(x * TWO_OVER_PI);
    double r = fma(-q, PIO2_HI, x);
    r = fma(-q, PIO2_LO, r);
    const double r2 = r * r;
    // sin(r) = r (1 - r^2/3! + r^4/5! - ...), cos(r) = 1 - r^2/2! + r^4/4! - ...
    double ps = -1.0 / 355687428096000.0;            // -1/17!
    ps = fma(ps, r2, scoef(1.0 / 1307674368000.0));          // 1/15!
    ps = fma(ps, r2, scoef(-1.0 / 6227020800.0));            // -1/13!
    ps = fma(ps, r2, scoef(1.0 / 39916800.0));               // 1/11!
    ps = fma(ps, r2, scoef(-1.0 / 362880.0));                // -1/9!
    ps = fma(ps, r2, scoef(1.0 / 5040.0));                   // 1/7!
    ps = fma(ps, r2, scoef(-1.0 / 120.0));                   // -1/5!
    ps = fma(ps, r2, scoef(1.0 / 6.0));                      // 1/3! (sign applied below)
    const double sr = fma(-r * r2, ps, r);
    double pc = 1.0 / 6402373705728000.0;             // 1/18!
    pc = fma(pc, r2, scoef(-1.0 / 20922789888000.0));        // -1/16!
    pc = fma(pc, r2, scoef(1.0 / 87178291200.0));            // 1/14!
    pc = fma(pc, r2, scoef(-1.0 / 479001600.0));             // -1/12!
    pc = fma(pc, r2, scoef(1.0 / 3628800.0));                // 1/10!
    pc = fma(pc, r2, scoef(-1.0 / 40320.0));                 // -1/8!
    pc = fma(pc, r2, scoef(1.0 / 720.0));                    // 1/6!
    pc = fma(pc, r2, scoef(-1.0 / 24.0));                    // -1/4!
    pc = fma(pc, r2, 0.5);                                   // 1/2! (sign applied below)
    const double cr = fma(-r2, pc, 1.0);
    const int k = ((int)q) & 3;
    const double ss = (k & 1) ? cr : sr;
    const double cc = (k & 1) ? sr : cr;
    *s = (k & 2) ? -ss : ss;
    *c = ((k + 1) & 2) ? -cc : cc;
}

// e^{i theta} for large |theta| by table: theta reduced modulo 2 pi, then
// r = j pi/64 + b with integer |j| <= 64 and |b| <= pi/128; e^{i j pi/64} from a
// 129-entry table (built by fill_etab, normally in LDS) and e^{i b} from its
// series to b^7 / b^6 (truncation < 1e-18 relative): ~24 operations against
// ~35 for sincos_small after reduce_2pi, within a few ulp of it. Any theta (also NaN or inf)
// gives an index in [0, 128].
constexpr double PI64_HI = 0.04908738521234052;     // pi / 64 rounded
constexpr double PI64_LO = 1.9135106236677393e-18;  // pi / 64 - PI64_HI
template <class E>
__device__ inline cd cexpi_tab(double theta, const E* __restrict__ tab)
{
    const double r = reduce_2pi(theta);
    const double j = rint(r * 20.371832715762604);  // 64 / pi
    double b = fma(-j, PI64_HI, r);
    b = fma(-j, PI64_LO, b);
    const E e = tab[(int)j + 64];
    const double b2 = b * b;
    double ps = fma(b2, -1.0 / 5040.0, 1.0 / 120.0);
    ps = fma(b2, ps, -1.0 / 6.0);
    const double sb = fma(b * b2, ps, b);
    double pc = fma(b2, -1.0 / 720.0, 1.0 / 24.0);
    pc = fma(b2, pc, -0.5);
    const double cb = fma(b2, pc, 1.0);
    return cd{e.x * cb - e.y * sb, e.y * cb + e.x * sb};
}

// The table of cexpi_tab: tab[j] = (cos, sin)((j - 64) pi / 64), j = 0..128,
// filled by threads tid = 0..nt-1 of a workgroup (a barrier must follow).
__device__ inline void fill_etab(double2* tab, int tid, int nt)
{
    for (int j = tid; j < 129; j += nt) {
        double sn, cs;
        sincos_small(fma((double)(j - 64), PI64_HI, (double)(j - 64) * PI64_LO), &sn, &cs);
        tab[j] = make_double2(cs, sn);
    }
}

// e^{i j pi/64}, j = -64..64 (index j + 64), correctly rounded: the table of
// cexpi_tab in constant memory for code without an LDS copy.
struct ETabEntry {
    double x, y;
};
__constant__ const ETabEntry ETAB[129] = {
    {-1.0, 0.0},
    {-0.9987954562051724, -0.049067674327418015},
    {-0.9951847266721969, -0.0980171403295606},
    {-0.989176509964781, -0.14673047445536175},
    {-0.9807852804032304, -0.19509032201612828},
    {-0.970031253194544, -0.2429801799032639},
    {-0.9569403357322088, -0.2902846772544624},
    {-0.9415440651830208, -0.33688985339222005},
    {-0.9238795325112867, -0.3826834323650898},
    {-0.9039892931234433, -0.4275550934302821},
    {-0.881921264348355, -0.47139673682599764},
    {-0.8577286100002721, -0.5141027441932218},
    {-0.8314696123025452, -0.5555702330196022},
    {-0.8032075314806449, -0.5956993044924334},
    {-0.773010453362737, -0.6343932841636455},
    {-0.7409511253549591, -0.6715589548470184},
    {-0.7071067811865476, -0.7071067811865476},
    {-0.6715589548470184, -0.7409511253549591},
    {-0.6343932841636455, -0.773010453362737},
    {-0.5956993044924334, -0.8032075314806449},
    {-0.5555702330196022, -0.8314696123025452},
    {-0.5141027441932218, -0.8577286100002721},
    {-0.47139673682599764, -0.881921264348355},
    {-0.4275550934302821, -0.9039892931234433},
    {-0.3826834323650898, -0.9238795325112867},
    {-0.33688985339222005, -0.9415440651830208},
    {-0.2902846772544624, -0.9569403357322088},
    {-0.2429801799032639, -0.970031253194544},
    {-0.19509032201612828, -0.9807852804032304},
    {-0.14673047445536175, -0.989176509964781},
    {-0.0980171403295606, -0.9951847266721969},
    {-0.049067674327418015, -0.9987954562051724},
    {-1.891501989729776e-58, -1.0},
    {0.049067674327418015, -0.9987954562051724},
    {0.0980171403295606, -0.9951847266721969},
    {0.14673047445536175, -0.989176509964781},
    {0.19509032201612828, -0.9807852804032304},
    {0.2429801799032639, -0.970031253194544},
    {0.2902846772544624, -0.9569403357322088},
    {0.33688985339222005, -0.9415440651830208},
    {0.3826834323650898, -0.9238795325112867},
    {0.4275550934302821, -0.9039892931234433},
    {0.47139673682599764, -0.881921264348355},
    {0.5141027441932218, -0.8577286100002721},
    {0.5555702330196022, -0.8314696123025452},
    {0.5956993044924334, -0.8032075314806449},
    {0.6343932841636455, -0.773010453362737},
    {0.6715589548470184, -0.7409511253549591},
    {0.7071067811865476, -0.7071067811865476},
    {0.7409511253549591, -0.6715589548470184},
    {0.773010453362737, -0.6343932841636455},
    {0.8032075314806449, -0.5956993044924334},
    {0.8314696123025452, -0.5555702330196022},
    {0.8577286100002721, -0.5141027441932218},
    {0.881921264348355, -0.47139673682599764},
    {0.9039892931234433, -0.4275550934302821},
    {0.9238795325112867, -0.3826834323650898},
    {0.9415440651830208, -0.33688985339222005},
    {0.9569403357322088, -0.2902846772544624},
    {0.970031253194544, -0.2429801799032639},
    {0.9807852804032304, -0.19509032201612828},
    {0.989176509964781, -0.14673047445536175},
    {0.9951847266721969, -0.0980171403295606},
    {0.9987954562051724, -0.049067674327418015},
    {1.0, 0.0},
    {0.9987954562051724, 0.049067674327418015},
    {0.9951847266721969, 0.0980171403295606},
    {0.989176509964781, 0.14673047445536175},
    {0.9807852804032304, 0.19509032201612828},
    {0.970031253194544, 0.2429801799032639},
    {0.9569403357322088, 0.2902846772544624},
    {0.9415440651830208, 0.33688985339222005},
    {0.9238795325112867, 0.3826834323650898},
    {0.9039892931234433, 0.4275550934302821},
    {0.881921264348355, 0.47139673682599764},
    {0.8577286100002721, 0.5141027441932218},
    {0.8314696123025452, 0.5555702330196022},
    {0.8032075314806449, 0.5956993044924334},
    {0.773010453362737, 0.6343932841636455},
    {0.7409511253549591, 0.6715589548470184},
    {0.7071067811865476, 0.7071067811865476},
    {0.6715589548470184, 0.7409511253549591},
    {0.6343932841636455, 0.773010453362737},
    {0.5956993044924334, 0.8032075314806449},
    {0.5555702330196022, 0.8314696123025452},
    {0.5141027441932218, 0.8577286100002721},
    {0.47139673682599764, 0.881921264348355},
    {0.4275550934302821, 0.9039892931234433},
    {0.3826834323650898, 0.9238795325112867},
    {0.33688985339222005, 0.9415440651830208},
    {0.2902846772544624, 0.9569403357322088},
    {0.2429801799032639, 0.970031253194544},
    {0.19509032201612828, 0.9807852804032304},
    {0.14673047445536175, 0.989176509964781},
    {0.0980171403295606, 0.9951847266721969},
    {0.049067674327418015, 0.9987954562051724},
    {-1.891501989729776e-58, 1.0},
    {-0.049067674327418015, 0.9987954562051724},
    {-0.0980171403295606, 0.9951847266721969},
    {-0.14673047445536175, 0.989176509964781},
    {-0.19509032201612828, 0.9807852804032304},
    {-0.2429801799032639, 0.970031253194544},
    {-0.2902846772544624, 0.9569403357322088},
    {-0.33688985339222005, 0.9415440651830208},
    {-0.3826834323650898, 0.9238795325112867},
    {-0.4275550934302821, 0.9039892931234433},
    {-0.47139673682599764, 0.881921264348355},
    {-0.5141027441932218, 0.8577286100002721},
    {-0.5555702330196022, 0.8314696123025452},
    {-0.5956993044924334, 0.8032075314806449},
    {-0.6343932841636455, 0.773010453362737},
    {-0.6715589548470184, 0.7409511253549591},
    {-0.7071067811865476, 0.7071067811865476},
    {-0.7409511253549591, 0.6715589548470184},
    {-0.773010453362737, 0.6343932841636455},
    {-0.8032075314806449, 0.5956993044924334},
    {-0.8314696123025452, 0.5555702330196022},
    {-0.8577286100002721, 0.5141027441932218},
    {-0.881921264348355, 0.47139673682599764},
    {-0.9039892931234433, 0.4275550934302821},
    {-0.9238795325112867, 0.3826834323650898},
    {-0.9415440651830208, 0.33688985339222005},
    {-0.9569403357322088, 0.2902846772544624},
    {-0.970031253194544, 0.2429801799032639},
    {-0.9807852804032304, 0.19509032201612828},
    {-0.989176509964781, 0.14673047445536175},
    {-0.9951847266721969, 0.0980171403295606},
    {-0.9987954562051724, 0.049067674327418015},
    {-1.0, 0.0},
};

// e^{i theta} and sin, cos of any moderate or large argument (phases of
// 1e5-1e8 rad) by cexpi_tab on the constant table
__device__ inline cd cexpi_red(double theta) { return cexpi_tab(theta, ETAB); }
__device__ inline void sincos_red(double theta, double* s, double* c)
{
    const cd e = cexpi_tab(theta, ETAB);
    *s = e.im;
    *c = e.re;
}
__device__ inline cd cexpi(double phase) { return cexpi_tab(phase, ETAB); }

// i * a
__host__ __device__ inline cd mul_i(cd a) { return cd{-a.im, a.re}; }

inline int ceil_div(long a, long b) { return int((a + b - 1) / b); }

}  // namespace bbhx
