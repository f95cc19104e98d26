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

// Number of distinct segments per mode that a wavefront reads with
// wavefront-uniform (scalar) loads before its remaining lanes load their own.
constexpr int INTERP_UNIFORM_SEGS = 3;

// One spline segment of one mode, repacked from the spline's arrays so that a
// wavefront reads it from one base address: c[p] = (y, c1, c2, c3) of quantity
// p (amp, phase, tf, T1re, T1im, T2re, T2im, T3re, T3im), x the segment's left
// knot. 320 B, five 64-B lines.
struct alignas(64) InterpSeg {
    double c[9][4];
    double x;
    double pad[3];
};

// Add segment r of a mode at dense frequency f to h (WaveformBuild.cu:48-60,
// :80-143): every quantity is the cubic y + c1 x + c2 x^2 + c3 x^3 in
// x = f - x_k; the mode counts only where t_start <= t(f) and
// (t(f) <= t_end or t_end <= 0).
template <class E>
__device__ inline void interp_add_segment(cd h[3], double f, const InterpSeg* __restrict__ r, const InterpBin& d,
                                          const E* tab)
{
    const double x = f - r->x;
    const double x2 = x * x, x3 = x * x2;
    double v[9];
#pragma unroll
    for (int p = 0; p < 9; p++) v[p] = r->c[p][0] + r->c[p][1] * x + r->c[p][2] * x2 + r->c[p][3] * x3;
    const double tf = v[2];
    if ((tf >= d.t_start) && ((tf <= d.t_end) || (d.t_end <= 0.0))) {
        const cd ap = v[0] * cexpi_tab(v[1], tab);
        h[0] += conj(mk(v[3], v[4]) * ap);
        h[1] += conj(mk(v[5], v[6]) * ap);
        h[2] += conj(mk(v[7], v[8]) * ap);
    }
}

// Sum of the modes at dense frequency f (bin i of the binary d whose segment
// records are R[(b numModes + m) length + k]):
//   h_ch(f) = sum_m conj(T_ch,m A_m e^{i phi_m}) on segment k = inds[m][i]
// (k <= -1 or k >= length - 1: the mode is left out). Dense bins are much finer
// than the sparse grid over most of a band, so the lanes of a wavefront usually
// share a segment: the first INTERP_UNIFORM_SEGS distinct segments of a mode
// are read once per wavefront at a wavefront-uniform address (scalar loads) and
// evaluated by the lanes that hold them; lanes left over (low frequencies,
// short segments, unsorted inds) read their own segment. Each lane adds one
// term per mode, modes in order, whichever way its segment was read.
template <class E>
__device__ inline void interp_point(cd h[3], double f, int i, const InterpBin& d, int b,
                                    const InterpSeg* __restrict__ R, int length, int numModes, const E* tab)
{
    h[0] = mk(0.0, 0.0);
    h[1] = mk(0.0, 0.0);
    h[2] = mk(0.0, 0.0);
    // each mode's segment index is loaded one mode ahead, so that its latency
    // overlaps the previous mode's work (the walker itself only waits on scalar
    // loads); through a global-address-space pointer, since a generic (flat) load
    // also counts against the scalar loads' counter and would be waited on with them
    const __attribute__((address_space(1))) int* inds = (const __attribute__((address_space(1))) int*)d.inds;
    int k_next = inds[i];
    // f resident before the loop: a first wait for it inside the loop would be
    // placed in every iteration and also wait for the prefetched index
    asm volatile("" ::"v"(f));
    for (int m = 0; m < numModes; m++) {
        const int k = k_next;
        if (m + 1 < numModes) k_next = inds[(long)(m + 1) * d.len + i];
        const long mbase = ((long)b * numModes + m) * length;
        bool pending = true;
#pragma unroll 1
        for (int it = 0; it < INTERP_UNIFORM_SEGS; it++) {
            const unsigned long long pm = __ballot(pending);
            if (pm == 0) break;
            const int ku = __builtin_amdgcn_readlane(k, __builtin_ctzll(pm));
            const bool mine = pending && k == ku;
            pending = pending && !mine;
            if ((ku <= -1) || (ku >= length - 1)) continue;
            // opaque, so that the compiler cannot substitute the lane's k (equal
            // to ku where 'mine' holds) and turn the scalar loads into vector ones
            long ub = mbase + ku;
            asm volatile("" : "+s"(ub));
            if (mine) interp_add_segment(h, f, R + ub, d, tab);
        }
        if (pending && (k > -1) && (k < length - 1)) interp_add_segment(h, f, R + mbase + k, d, tab);
    }
}

InterpBin* upload_table(const std::vector<InterpBin>& h, hipStream_t stream);
long build_table(std::vector<InterpBin>& h, const int64_t* tmpl_ptrs, const int64_t* inds_ptrs,
                 const int* inds_start, const int* ind_lengths, const double* t_start, const double* t_end,
                 int numBinAll, int tile);

}  // namespace bbhx
