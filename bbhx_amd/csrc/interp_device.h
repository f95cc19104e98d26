// Device helpers shared by the template-build and fused direct-likelihood kernels.
#pragma once

#include <vector>

#include "common.h"

namespace bbhx {

constexpr int INTERP_TILE = 256;

// Per-binary descriptor of a ragged dense band.
struct InterpBin {
    double* tmpl;      // (3, len) complex output (may be null for the fused likelihood)
    const int* inds;   // (M, len) int32 spline segment per mode and dense bin
    int start, len;    // band in the dense data grid
    double t_start, t_end;
    int tile_begin;    // exclusive prefix sum of tiles
    int pad;
};

// binary owning tile 'tile' (largest b with tile_begin <= tile)
__device__ inline int find_bin(const InterpBin* __restrict__ bins, int n, int tile)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (bins[mid].tile_begin <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Evaluate the 9 splines of every mode at dense frequency f (bin i of binary b)
// and combine (WaveformBuild.cu:48-60, :80-143).
__device__ inline void interp_point(cd h[3], double f, int i, const InterpBin& d, int b,
                                    const double* __restrict__ freqs, const double* __restrict__ y,
                                    const double* __restrict__ c1, const double* __restrict__ c2,
                                    const double* __restrict__ c3, int length, int numModes, long nper)
{
    h[0] = mk(0.0, 0.0);
    h[1] = mk(0.0, 0.0);
    h[2] = mk(0.0, 0.0);
    for (int m = 0; m < numModes; m++) {
        const int k = d.inds[(long)m * d.len + i];
        if ((k <= -1) || (k >= length - 1)) continue;
        const long base = ((long)b * numModes + m) * length + k;
        const double x = f - freqs[base];
        const double x2 = x * x, x3 = x * x2;
        double v[9];
#pragma unroll
        for (int p = 0; p < 9; p++) {
            const long q = p * nper + base;
            v[p] = y[q] + c1[q] * x + c2[q] * x2 + c3[q] * x3;
        }
        const double tf = v[2];
        if ((tf >= d.t_start) && ((tf <= d.t_end) || (d.t_end <= 0.0))) {
            const cd ap = v[0] * cexpi_red(v[1]);
            h[0] += conj(mk(v[3], v[4]) * ap);
            h[1] += conj(mk(v[5], v[6]) * ap);
            h[2] += conj(mk(v[7], v[8]) * ap);
        }
    }
}

InterpBin* upload_table(const std::vector<InterpBin>& h, hipStream_t stream);
long build_table(std::vector<InterpBin>& h, const int64_t* tmpl_ptrs, const int64_t* inds_ptrs,
                 const int* inds_start, const int* ind_lengths, const double* t_start, const double* t_end,
                 int numBinAll, int tile);

}  // namespace bbhx
