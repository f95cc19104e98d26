// FP64 CPU restatement of the batched not-a-knot cubic spline.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Follows cutils/Interpolate.cu of the reference: the same tridiagonal system
// for the knot slopes (Interpolate.cu:15-127, identical to scipy's
// CubicSpline with bc_type='not-a-knot'), solved the way the reference CPU
// path does, with Gaussian elimination and partial pivoting on the tridiagonal
// band (the algorithm of LAPACK dgtsv, which Interpolate.cu:236 calls; written
// here from its published description), then the polynomial coefficients of
// Interpolate.cu:245-261.

#include "oracle.h"

#include <cmath>
#include <utility>
#include <vector>

namespace {

// Solve a tridiagonal system in place with partial pivoting.
//   dl[1..n-1] sub-diagonal, d[0..n-1] diagonal, du[0..n-2] super-diagonal,
//   b[0..n-1] right-hand side -> solution.
void gtsv_pivot(int n, std::vector<double>& dl, std::vector<double>& d, std::vector<double>& du,
                double* b)
{
    std::vector<double> du2(n, 0.0);  // second super-diagonal created by row swaps
    for (int i = 0; i < n - 1; i++) {
        if (std::fabs(d[i]) >= std::fabs(dl[i + 1])) {
            // no interchange
            double mult = dl[i + 1] / d[i];
            d[i + 1] -= mult * du[i];
            b[i + 1] -= mult * b[i];
            dl[i + 1] = 0.0;
        } else {
            // interchange rows i and i+1
            double mult = d[i] / dl[i + 1];
            d[i] = dl[i + 1];
            double temp = d[i + 1];
            d[i + 1] = du[i] - mult * temp;
            if (i < n - 2) {
                du2[i] = du[i + 1];
                du[i + 1] = -mult * du2[i];
            }
            du[i] = temp;
            std::swap(b[i], b[i + 1]);
            b[i + 1] -= mult * b[i];
        }
    }
    // back substitution with the (up to) two super-diagonals
    b[n - 1] /= d[n - 1];
    if (n > 1) b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / d[n - 2];
    for (int i = n - 3; i >= 0; i--) b[i] = (b[i] - du[i] * b[i + 1] - du2[i] * b[i + 2]) / d[i];
}

}  // namespace

extern "C" {

// interpolate -> fill_B / prep_splines, interpolate_kern, set_spline_constants,
// Interpolate.cu:15-127, :201-241, :245-318, :320-359.
// x: (B, M, L); y: (P, B, M, L); outputs B (knot slopes), c1, c2, c3: (P, B, M, L).
void orc_interpolate(const double* x, const double* y, double* Bout, double* c1, double* c2,
                     double* c3, int length, int numInterpParams, int numModes, int numBinAll)
{
    const int nsub = numModes * numBinAll;
    const int ninterps = nsub * numInterpParams;
    const int n = length;

#pragma omp parallel for schedule(static)
    for (int s = 0; s < ninterps; s++) {
        const int sub = s % nsub;
        const double* xs = x + (long)sub * n;
        const double* ys = y + (long)s * n;
        double* slope = Bout + (long)s * n;
        std::vector<double> dl(n, 0.0), d(n, 0.0), du(n, 0.0);

        // row 0: not-a-knot condition at the left end
        {
            double dx1 = xs[1] - xs[0], dx2 = xs[2] - xs[1], dd = xs[2] - xs[0];
            double s1 = (ys[1] - ys[0]) / dx1, s2 = (ys[2] - ys[1]) / dx2;
            slope[0] = ((dx1 + 2 * dd) * dx2 * s1 + dx1 * dx1 * s2) / dd;
            d[0] = dx2;
            du[0] = dd;
        }
        for (int i = 1; i < n - 1; i++) {
            double dx1 = xs[i] - xs[i - 1], dx2 = xs[i + 1] - xs[i];
            double s1 = (ys[i] - ys[i - 1]) / dx1, s2 = (ys[i + 1] - ys[i]) / dx2;
            slope[i] = 3.0 * (dx2 * s1 + dx1 * s2);
            d[i] = 2 * (dx1 + dx2);
            du[i] = dx1;
            dl[i] = dx2;
        }
        // row n-1: not-a-knot condition at the right end
        {
            double dx1 = xs[n - 2] - xs[n - 3], dx2 = xs[n - 1] - xs[n - 2], dd = xs[n - 1] - xs[n - 3];
            double s1 = (ys[n - 2] - ys[n - 3]) / dx1, s2 = (ys[n - 1] - ys[n - 2]) / dx2;
            slope[n - 1] = ((dx2 * dx2 * s1 + (2 * dd + dx2) * dx1 * s2) / dd);
            d[n - 1] = dx1;
            dl[n - 1] = dd;
        }
        gtsv_pivot(n, dl, d, du, slope);

        for (int i = 0; i < n - 1; i++) {
            double dx = xs[i + 1] - xs[i];
            double sl = (ys[i + 1] - ys[i]) / dx;
            double t = (slope[i] + slope[i + 1] - 2 * sl) / dx;
            c1[(long)s * n + i] = slope[i];
            c2[(long)s * n + i] = (sl - slope[i]) / dx - t;
            c3[(long)s * n + i] = t / dx;
        }
        c1[(long)s * n + n - 1] = 0.0;
        c2[(long)s * n + n - 1] = 0.0;
        c3[(long)s * n + n - 1] = 0.0;
    }
}

}  // extern "C"
