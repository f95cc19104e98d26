// Batched not-a-knot cubic spline (the reference's CubicSplineInterpolant,
// Interpolate.cu:15-359, which solves with cuSPARSE gtsv2StridedBatch).
//
// Two solvers of the same tridiagonal systems:
//  * k_spline_part, one wavefront per system, for batches too small to fill
//    the chip with one lane per system (the heterodyne reference's 54 systems
//    of 8096 knots, a few templates): the rows are cut into up to 64
//    contiguous blocks, one per lane; each lane solves its block's interior
//    with the block's two end unknowns as boundary values (three Thomas right
//    sides: the data and the two unit boundary responses), which turns each
//    block's first and last rows into two rows of a reduced tridiagonal
//    system of 2 x (blocks) unknowns in LDS; one lane solves that by Thomas,
//    and every lane then forms its rows' unknowns and segment coefficients.
//    The dependency chain is ~3 L / 64 + 128 rows instead of 2 L.
//  * k_spline, one lane per system (parameter, binary, mode), for batches of
//    SPL_PART_MAX_SYSTEMS systems and more:
//
// The system for the knot
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

// ------------------------------------------------------------ partitioned solver

constexpr int PART_LANES = 64;

// Row i of the slope system (the rows k_spline builds): a s_{i-1} + b s_i + c s_{i+1} = d.
struct Row {
    double a, b, c, d;
};

__device__ inline Row spline_row(const double* __restrict__ xs, const double* __restrict__ ys, int i, int n)
{
    Row r;
    if (i == 0) {  // not-a-knot, reduced to two terms: dx2 s0 + (x2 - x0) s1 = b
        const double x0 = xs[0], x1 = xs[1], x2 = xs[2], y0 = ys[0], y1 = ys[1], y2 = ys[2];
        const double dx1 = x1 - x0, dx2 = x2 - x1, dd = x2 - x0;
        const double s1 = (y1 - y0) / dx1, s2 = (y2 - y1) / dx2;
        r.a = 0.0;
        r.b = dx2;
        r.c = dd;
        r.d = ((dx1 + 2 * dd) * dx2 * s1 + dx1 * dx1 * s2) / dd;
        return r;
    }
    if (i == n - 1) {  // not-a-knot: (x_{n-1} - x_{n-3}) s_{n-2} + (x_{n-2} - x_{n-3}) s_{n-1} = rhs
        const double xm3 = xs[n - 3], xm2 = xs[n - 2], xm1 = xs[n - 1];
        const double ym3 = ys[n - 3], ym2 = ys[n - 2], ym1 = ys[n - 1];
        const double ex1 = xm2 - xm3, ex2 = xm1 - xm2, ed = xm1 - xm3;
        const double t1 = (ym2 - ym3) / ex1, t2 = (ym1 - ym2) / ex2;
        r.a = ed;
        r.b = ex1;
        r.c = 0.0;
        r.d = (ex2 * ex2 * t1 + (2 * ed + ex2) * ex1 * t2) / ed;
        return r;
    }
    const double xa = xs[i - 1], xb = xs[i], xc = xs[i + 1], ya = ys[i - 1], yb = ys[i], yc = ys[i + 1];
    const double ex1 = xb - xa, ex2 = xc - xb;
    const double t1 = (yb - ya) / ex1, t2 = (yc - yb) / ex2;
    r.a = ex2;
    r.b = 2 * (ex1 + ex2);
    r.c = ex1;
    r.d = 3.0 * (ex2 * t1 + ex1 * t2);
    return r;
}

// One wavefront per system s0 + blockIdx.x. Lane j < P owns rows [st, en] of
// P = min(64, n / 3) contiguous blocks of at least 3 rows. Scratch: the four
// output slabs of the system (per row: c' -> c3, then the three solutions
// p -> B, q -> c1, r -> c2 of the block interior with the end unknowns z_st,
// z_en as boundary values, z_i = p_i + q_i z_st + r_i z_en).
__global__ __launch_bounds__(PART_LANES) void k_spline_part(const double* __restrict__ x,
                                                            const double* __restrict__ y, double* __restrict__ Bs,
                                                            double* __restrict__ c1, double* __restrict__ c2,
                                                            double* __restrict__ c3, int n, int nsub, int s0)
{
    __shared__ double red[4][2 * PART_LANES];  // reduced rows: a, b, c, d
    __shared__ double zr[2 * PART_LANES];
    const int s = s0 + (int)blockIdx.x;
    const int j = threadIdx.x;
    const int P = n / 3 < PART_LANES ? n / 3 : PART_LANES;
    const double* xs = x + (long)(s % nsub) * n;
    const double* ys = y + (long)s * n;
    double* pB = Bs + (long)s * n;
    double* q1 = c1 + (long)s * n;
    double* r2 = c2 + (long)s * n;
    double* cp3 = c3 + (long)s * n;
    const int base = n / P, rem = n % P;
    const int st = j * base + (j < rem ? j : rem);
    const int en = st + base + (j < rem ? 1 : 0) - 1;
    if (j < P) {
        // interior rows st+1 .. en-1, forward sweep of the three right sides
        double cprev = 0.0, fp = 0.0, fq = 0.0, fr = 0.0;
        for (int i = st + 1; i <= en - 1; i++) {
            const Row w = spline_row(xs, ys, i, n);
            const double im = 1.0 / (w.b - w.a * cprev);
            const double rq = (i == st + 1) ? -w.a : 0.0;  // a z_st moved to the right side
            const double rr = (i == en - 1) ? -w.c : 0.0;  // c z_en moved to the right side
            cprev = w.c * im;
            fp = (w.d - w.a * fp) * im;
            fq = (rq - w.a * fq) * im;
            fr = (rr - w.a * fr) * im;
            cp3[i] = cprev;
            pB[i] = fp;
            q1[i] = fq;
            r2[i] = fr;
        }
        // backward sweep: p, q, r in place
        double pn = 0.0, qn = 0.0, rn = 0.0;
        for (int i = en - 1; i >= st + 1; i--) {
            const double cc = cp3[i];
            pn = pB[i] - cc * pn;
            qn = q1[i] - cc * qn;
            rn = r2[i] - cc * rn;
            pB[i] = pn;
            q1[i] = qn;
            r2[i] = rn;
        }
        // rows st and en in the unknowns (z_en of the block before, z_st, z_en) and
        // (z_st, z_en, z_st of the block after): rows 2j and 2j + 1 of the reduced system
        const Row w0 = spline_row(xs, ys, st, n);
        const double p1 = pB[st + 1], q1v = q1[st + 1], r1v = r2[st + 1];
        red[0][2 * j] = w0.a;
        red[1][2 * j] = w0.b + w0.c * q1v;
        red[2][2 * j] = w0.c * r1v;
        red[3][2 * j] = w0.d - w0.c * p1;
        const Row w1 = spline_row(xs, ys, en, n);
        const double pm = pB[en - 1], qm = q1[en - 1], rm = r2[en - 1];
        red[0][2 * j + 1] = w1.a * qm;
        red[1][2 * j + 1] = w1.b + w1.a * rm;
        red[2][2 * j + 1] = w1.c;
        red[3][2 * j + 1] = w1.d - w1.a * pm;
    }
    __syncthreads();
    if (j == 0) {  // the reduced system (2P rows) by Thomas
        const int m = 2 * P;
        double cprev = 0.0, dprev = 0.0;
        for (int i = 0; i < m; i++) {
            const double im = 1.0 / (red[1][i] - red[0][i] * cprev);
            cprev = red[2][i] * im;
            dprev = (red[3][i] - red[0][i] * dprev) * im;
            red[2][i] = cprev;
            red[3][i] = dprev;
        }
        double z = red[3][m - 1];
        zr[m - 1] = z;
        for (int i = m - 2; i >= 0; i--) {
            z = red[3][i] - red[2][i] * z;
            zr[i] = z;
        }
    }
    __syncthreads();
    if (j < P) {
        // slopes of the block and the segment coefficients (k_spline's backward
        // sweep formulas): c1 = s_i, c2 = (m_i - s_i)/dx - t, c3 = t/dx,
        // t = (s_i + s_{i+1} - 2 m_i)/dx; the last knot's entries are 0
        const double zs = zr[2 * j], ze = zr[2 * j + 1];
        double zi = zs;
        for (int i = st; i <= en; i++) {
            double znext;
            if (i + 1 <= en - 1) znext = pB[i + 1] + q1[i + 1] * zs + r2[i + 1] * ze;
            else if (i + 1 == en) znext = ze;
            else znext = (j + 1 < P) ? zr[2 * j + 2] : 0.0;
            pB[i] = zi;
            if (i == n - 1) {
                q1[i] = 0.0;
                r2[i] = 0.0;
                cp3[i] = 0.0;
            } else {
                const double xi = xs[i], xn = xs[i + 1];
                const double idx = 1.0 / (xn - xi);
                const double slope = (ys[i + 1] - ys[i]) * idx;
                const double t = (zi + znext - 2 * slope) * idx;
                q1[i] = zi;
                r2[i] = (slope - zi) * idx - t;
                cp3[i] = t * idx;
            }
            zi = znext;
        }
    }
}

// Systems below this count go to the partitioned solver (one wavefront each).
// Measured on MI355X (9 quantities x binaries x 6 modes): 54 systems of 8096
// knots 4.02 -> 0.21 ms, 864 of 1024 knots 0.55 -> 0.16 ms; but 3456 of 4096
// 2.34 -> 3.93 ms and 27648 of 4096 6.5 -> 33 ms, where its per-lane row blocks
// (64 cache lines per load) and scratch round trips cost more than the
// one-lane-per-system sweeps, which by then no longer wait on one chain.
constexpr long SPL_PART_MAX_SYSTEMS = 2048;

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
    const bool part = nsys < SPL_PART_MAX_SYSTEMS;
    for (long s0 = 0; s0 < nsys; s0 += per) {
        const long ns = nsys - s0 < per ? nsys - s0 : per;
        if (part)
            hipLaunchKernelGGL(k_spline_part, dim3((unsigned)ns), dim3(PART_LANES), 0,
                               reinterpret_cast<hipStream_t>(stream), freqs, propArrays, B, upper_diag, diag,
                               lower_diag, length, numModes * numBinAll, (int)s0);
        else
            hipLaunchKernelGGL(k_spline, dim3(ceil_div(ns, 64)), dim3(64), 0, reinterpret_cast<hipStream_t>(stream),
                               freqs, propArrays, B, upper_diag, diag, lower_diag, length, numModes * numBinAll,
                               (int)s0, (int)nsys);
        if (hipGetLastError() != hipSuccess) return ERR_LAUNCH;
    }
    return OK;
}
