// Batched not-a-knot cubic spline (the reference's CubicSplineInterpolant,
// Interpolate.cu:15-359, which solves with cuSPARSE gtsv2StridedBatch).
//
// One lane per system (parameter, binary, mode). The system for the knot
// slopes s_i is the reference's (identical to scipy CubicSpline,
// bc_type='not-a-knot'): interior rows dx_i s_{i-1} + 2(dx_{i-1}+dx_i) s_i +
// dx_{i-1} s_{i+1} = 3(dx_i m_{i-1} + dx_{i-1} m_i), with the not-a-knot end
// rows. It is solved by the Thomas algorithm (LU without pivoting; the system
// is nonsingular and the pivots stay positive for increasing knots, see the
// end-row analysis in the test suite), building each row on the fly from x and
// y, so no band matrix is ever materialised. A lane's sweeps are one long
// dependent chain: each sweep loads SPL_BLOCK rows ahead of the recurrence
// (one memory latency per block) and divides once per row (a reciprocal
// instead of two or four divisions):
//   forward sweep: c'_i -> lower_diag slab, d'_i -> B slab
//   backward sweep: s_i -> B, then the segment coefficients
//   c1 = s_i, c2 = (m_i - s_i)/dx - t, c3 = t/dx, t = (s_i + s_{i+1} - 2 m_i)/dx
// written to upper_diag (c1), diag (c2), lower_diag (c3) as in the reference
// (Interpolate.cu:245-261, :332-334). The last knot's c1..c3 entries are set to 0.

#include "bbhx_hip.h"
#include "common.h"

using namespace bbhx;

namespace {

constexpr int SPL_BLOCK = 16;  // rows whose loads are issued together in each sweep

// Systems per launch are capped so that the arrays the launch's lanes walk
// through at once (x, y and the three outputs, 40 B per row and system) stay
// within about 2.5 GiB: each lane strides through its own rows, so the launch
// touches all of them concurrently, and beyond that size the sweeps slow down
// more than in proportion to the work (measured: 256 x 6 x 9 systems of 4096
// knots 2.9 ms, twice as many 7.3 ms, four times as many 28 ms in one launch).
constexpr double SPL_FOOTPRINT_CAP = 2.5 * (1L << 30);

__global__ __launch_bounds__(64) void k_spline(const double* __restrict__ x, const double* __restrict__ y,
                                               double* __restrict__ Bs, double* __restrict__ c1,
                                               double* __restrict__ c2, double* __restrict__ c3, int n, int nsub,
                                               int s0, int nsys)
{
    const int s = s0 + (int)blockIdx.x * 64 + (int)threadIdx.x;
    if (s >= nsys) return;
    const int sub = s % nsub;
    const double* xs = x + (long)sub * n;
    const double* ys = y + (long)s * n;
    double* dp = Bs + (long)s * n;  // d' then the slopes
    double* cp = c3 + (long)s * n;  // c' then c3

    // row 0 (not-a-knot)
    double x0 = xs[0], x1 = xs[1], x2 = xs[2];
    double y0 = ys[0], y1 = ys[1], y2 = ys[2];
    {
        double dx1 = x1 - x0, dx2 = x2 - x1, dd = x2 - x0;
        double s1 = (y1 - y0) / dx1, s2 = (y2 - y1) / dx2;
        double b = ((dx1 + 2 * dd) * dx2 * s1 + dx1 * dx1 * s2) / dd;
        double cprev = dd / dx2;
        double dprev = b / dx2;
        cp[0] = cprev;
        dp[0] = dprev;
        // interior rows, in blocks of SPL_BLOCK whose loads are all issued before
        // the recurrence runs through the block (one memory latency per block,
        // not per row)
        double xa = x0, xb = x1, ya = y0, yb = y1;
        for (int i0 = 1; i0 < n - 1; i0 += SPL_BLOCK) {
            double xv[SPL_BLOCK], yv[SPL_BLOCK];
#pragma unroll
            for (int u = 0; u < SPL_BLOCK; u++) {
                const int i = i0 + u;
                xv[u] = i < n - 1 ? xs[i + 1] : 0.0;
                yv[u] = i < n - 1 ? ys[i + 1] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < SPL_BLOCK; u++) {
                const int i = i0 + u;
                if (i >= n - 1) break;
                const double xc = xv[u], yc = yv[u];
                double ex1 = xb - xa, ex2 = xc - xb;
                double t1 = (yb - ya) / ex1, t2 = (yc - yb) / ex2;
                double rhs = 3.0 * (ex2 * t1 + ex1 * t2);
                double diag = 2 * (ex1 + ex2), lo = ex2, up = ex1;
                const double im = 1.0 / (diag - lo * cprev);
                cprev = up * im;
                dprev = (rhs - lo * dprev) * im;
                cp[i] = cprev;
                dp[i] = dprev;
                xa = xb; xb = xc; ya = yb; yb = yc;
            }
        }
        // row n-1 (not-a-knot)
        double xm3 = xs[n - 3], xm2 = xs[n - 2], xm1 = xs[n - 1];
        double ym3 = ys[n - 3], ym2 = ys[n - 2], ym1 = ys[n - 1];
        double ex1 = xm2 - xm3, ex2 = xm1 - xm2, ed = xm1 - xm3;
        double t1 = (ym2 - ym3) / ex1, t2 = (ym1 - ym2) / ex2;
        double rhs = ((ex2 * ex2 * t1 + (2 * ed + ex2) * ex1 * t2) / ed);
        double m = ex1 - ed * cprev;
        dp[n - 1] = (rhs - ed * dprev) / m;
    }

    // backward sweep fused with the segment coefficients
    double snext = dp[n - 1];
    double xn = xs[n - 1], yn = ys[n - 1];
    c1[(long)s * n + n - 1] = 0.0;
    c2[(long)s * n + n - 1] = 0.0;
    cp[n - 1] = 0.0;
    for (int i0 = n - 2; i0 >= 0; i0 -= SPL_BLOCK) {
        // rows i0, i0 - 1, ...: every load of the block before its first store
        double dv[SPL_BLOCK], cv[SPL_BLOCK], xv[SPL_BLOCK], yv[SPL_BLOCK];
#pragma unroll
        for (int u = 0; u < SPL_BLOCK; u++) {
            const int i = i0 - u;
            dv[u] = i >= 0 ? dp[i] : 0.0;
            cv[u] = i >= 0 ? cp[i] : 0.0;
            xv[u] = i >= 0 ? xs[i] : 0.0;
            yv[u] = i >= 0 ? ys[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < SPL_BLOCK; u++) {
            const int i = i0 - u;
            if (i < 0) break;
            double si = dv[u] - cv[u] * snext;
            dp[i] = si;
            double xi = xv[u], yi = yv[u];
            const double idx = 1.0 / (xn - xi);
            const double slope = (yn - yi) * idx;
            const double t = (si + snext - 2 * slope) * idx;
            c1[(long)s * n + i] = si;
            c2[(long)s * n + i] = (slope - si) * idx - t;
            cp[i] = t * idx;
            snext = si;
            xn = xi;
            yn = yi;
        }
    }
}

}  // namespace

extern "C" int bbhx_hip_interpolate(const double* freqs, const double* propArrays, double* B, double* upper_diag,
                                    double* diag, double* lower_diag, int length, int numInterpParams, int numModes,
                                    int numBinAll, void* stream)
{
    if (length < 3 || numInterpParams < 1 || numModes < 1 || numBinAll < 0) return ERR_ARGS;
    if (numBinAll == 0) return OK;
    if (!freqs || !propArrays || !B || !upper_diag || !diag || !lower_diag) return ERR_ARGS;
    long nsys = (long)numInterpParams * numModes * numBinAll;
    if (nsys > 2147483647L) return ERR_ARGS;
    // launches of whole wavefronts of systems, each within the footprint cap
    long per = (long)(SPL_FOOTPRINT_CAP / (40.0 * length)) / 64 * 64;
    if (per < 64) per = 64;
    for (long s0 = 0; s0 < nsys; s0 += per) {
        const long ns = nsys - s0 < per ? nsys - s0 : per;
        hipLaunchKernelGGL(k_spline, dim3(ceil_div(ns, 64)), dim3(64), 0, reinterpret_cast<hipStream_t>(stream),
                           freqs, propArrays, B, upper_diag, diag, lower_diag, length, numModes * numBinAll, (int)s0,
                           (int)nsys);
        if (hipGetLastError() != hipSuccess) return ERR_LAUNCH;
    }
    return OK;
}
