// Heterodyne (relative-binning) constants of a reference template
// (HeterodynedLikelihood.init_heterodyne_info, likelihood.py:499-533; the
// reference also has a kernel for it, prep_hdyn, Likelihood.cu:238-352, with
// double atomics, that its Python does not call). For sparse nodes f_0 < ... < f_{K-1} and the
// dense points k of segment j (f_j <= f_k < f_{j+1}), with
// w_k = 4 df / S_k:
//   A0[j+1] = sum w conj(h0) d,      A1[j+1] = sum w conj(h0) d (f - fm_j),
//   B0[j+1] = sum w |h0|^2,          B1[j+1] = sum w |h0|^2 (f - fm_j),
// per channel, fm_j the segment midpoint; the last dense point of the band is
// left out of the constants (the reference loops over unique(bins[:-1])), slot
// 0 stays zero. Also the totals d_d, h_h and Re d_h over every dense point.
//
// Deterministic: the dense points of each segment are cut into work items of
// at most PREP_CHUNK points; a persistent kernel reduces each item in a fixed
// order; a last kernel adds the items of each segment in item order.
//   k_prep_bounds   one workgroup: segment starts (binary search), item counts,
//                   exclusive scan
//   k_prep_items    persistent: one item per workgroup pass, 27 partial sums
//   k_prep_reduce   one lane per (slot, channel, constant)
//   k_prep_totals   one wavefront per total

#include "bbhx_hip.h"
#include "common.h"
#include "workspace.h"

using namespace bbhx;

namespace {

constexpr int PREP_CHUNK = 2048;
constexpr int NP = 27;  // per item: 4 constants x 3 channels x (re, im), then d_d, h_h, Re d_h

// first index k in [0, n) with a[k] >= v (n if none)
__device__ inline int lower_bound(const double* __restrict__ a, int n, double v)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// seg[j] = first dense point of segment j (j < K), seg[K] = N; item[j] = first
// work item of segment j, item[K] = number of items.
__global__ __launch_bounds__(1024) void k_prep_bounds(int* __restrict__ seg, int* __restrict__ item,
                                                      const double* __restrict__ f, int N,
                                                      const double* __restrict__ nodes, int K)
{
    __shared__ int part[1024];
    const int t = threadIdx.x;
    for (int j = t; j <= K; j += 1024) seg[j] = j < K ? lower_bound(f, N, nodes[j]) : N;
    __syncthreads();
    const int per = (K + 1023) / 1024;
    const int j0 = t * per, j1 = min(K, j0 + per);
    int sum = 0;
    for (int j = j0; j < j1; j++) {
        const int n = seg[j + 1] - seg[j];
        const int c = n > 0 ? (n + PREP_CHUNK - 1) / PREP_CHUNK : 0;
        item[j] = c;  // count for now
        sum += c;
    }
    part[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = t > 0 ? part[t - 1] : 0;
    for (int j = j0; j < j1; j++) {
        const int c = item[j];
        item[j] = run;
        run += c;
    }
    if (t == 1023) item[K] = part[1023];
}

// segment owning work item it (largest j with item[j] <= it)
__device__ inline int item_segment(const int* __restrict__ item, int K, int it)
{
    int lo = 0, hi = K - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (item[mid] <= it) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ inline double wsum64(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(256) void k_prep_items(double* __restrict__ partial, const int* __restrict__ seg,
                                                    const int* __restrict__ item, const double* __restrict__ f,
                                                    const cd* __restrict__ d, const cd* __restrict__ h0,
                                                    const double* __restrict__ S, int N,
                                                    const double* __restrict__ nodes, int K, double df)
{
    __shared__ double red[NP][4];
    const int nitems = item[K];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int j = item_segment(item, K, it);
        const int a = seg[j] + (it - item[j]) * PREP_CHUNK;
        const int e = min(seg[j + 1], a + PREP_CHUNK);
        // segment midpoint (the last segment j = K - 1 feeds only the totals)
        const double fm = j < K - 1 ? (nodes[j] + nodes[j + 1]) / 2 : 0.0;
        double acc[NP];
#pragma unroll
        for (int p = 0; p < NP; p++) acc[p] = 0.0;
        for (int k = a + tid; k < e; k += blockDim.x) {
            const double dfk = f[k] - fm;
            const bool in_consts = (k < N - 1) && (j < K - 1);
#pragma unroll
            for (int ch = 0; ch < 3; ch++) {
                const long q = (long)ch * N + k;
                const double w = 4.0 / S[q] * df;
                const cd hv = h0[q], dv = d[q];
                // conj(h0) d, |h0|^2, |d|^2
                const double cre = hv.re * dv.re + hv.im * dv.im, cim = hv.re * dv.im - hv.im * dv.re;
                const double hh = hv.re * hv.re + hv.im * hv.im;
                const double dd = dv.re * dv.re + dv.im * dv.im;
                const double a0r = w * cre, a0i = w * cim, b0 = w * hh;
                if (in_consts) {
                    acc[(0 * 3 + ch) * 2] += a0r;
                    acc[(0 * 3 + ch) * 2 + 1] += a0i;
                    acc[(1 * 3 + ch) * 2] += a0r * dfk;
                    acc[(1 * 3 + ch) * 2 + 1] += a0i * dfk;
                    acc[(2 * 3 + ch) * 2] += b0;
                    acc[(3 * 3 + ch) * 2] += b0 * dfk;
                }
                acc[24] += w * dd;
                acc[25] += b0;
                acc[26] += a0r;
            }
        }
#pragma unroll
        for (int p = 0; p < NP; p++) {
            const double v = wsum64(acc[p]);
            if (lane == 0) red[p][wv] = v;
        }
        __syncthreads();
        if (tid < NP) {
            double s = 0.0;
            for (int w = 0; w < nw; w++) s += red[tid][w];
            partial[(long)it * NP + tid] = s;
        }
        __syncthreads();
    }
}

// lanes [0, 24 K): constant slot s = lane / 24, component p = lane % 24 (slot 0 and
// the unused imaginary parts of B0, B1 stay zero); a segment holds a few items.
__global__ void k_prep_reduce(double* __restrict__ consts, const double* __restrict__ partial,
                              const int* __restrict__ item, int K)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 24 * K) return;
    const int slot = t / 24, p = t % 24;
    const int c = p >> 1, ri = p & 1;  // c = constant * 3 + channel
    double s = 0.0;
    if (slot > 0) {
        const int j = slot - 1;
        for (int it = item[j]; it < item[j + 1]; it++) s += partial[(long)it * NP + p];
    }
    consts[2 * ((long)c * K + slot) + ri] = s;
}

// The three totals (d_d, h_h, Re d_h) sum every work item: one wavefront per
// total, lane l taking items l, l + 64, ... and the lanes combined by a fixed
// shuffle tree (deterministic; a single lane's dependent loop over the ~N/2048
// items took 140 us at N = 2^20).
__global__ __launch_bounds__(192) void k_prep_totals(double* __restrict__ totals, const double* __restrict__ partial,
                                                     const int* __restrict__ item, int K)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int p = 24 + w;
    double s = 0.0;
    for (int it = lane; it < item[K]; it += 64) s += partial[(long)it * NP + p];
    s = wsum64(s);
    if (lane == 0) totals[w] = s;
}

}  // namespace

extern "C" int bbhx_hip_hdyn_prep(double* dataConstants, double* totals, const double* dataFreqs,
                                  const double* dataChannels, const double* h0, const double* psd, int data_length,
                                  const double* nodes, int K, double df, void* stream)
{
    if (data_length < 2 || K < 2 || !(df > 0.0)) return ERR_ARGS;
    if (!dataConstants || !totals || !dataFreqs || !dataChannels || !h0 || !psd || !nodes) return ERR_ARGS;
    const long max_items = (long)(data_length + PREP_CHUNK - 1) / PREP_CHUNK + K;
    const size_t ints = 2 * ((size_t)K + 1);
    const size_t bytes = sizeof(double) * NP * (size_t)max_items + sizeof(int) * ints + 256;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    char* ws = static_cast<char*>(workspace_get(WS_PREP, bytes, st));
    if (!ws) return ERR_ALLOC;
    double* partial = reinterpret_cast<double*>(ws);
    int* seg = reinterpret_cast<int*>(ws + sizeof(double) * NP * (size_t)max_items);
    int* item = seg + (K + 1);
    hipLaunchKernelGGL(k_prep_bounds, dim3(1), dim3(1024), 0, st, seg, item, dataFreqs, data_length, nodes, K);
    // persistent: enough workgroups to fill the chip, each looping over items
    hipLaunchKernelGGL(k_prep_items, dim3(1024), dim3(256), 0, st, partial, seg, item, dataFreqs,
                       reinterpret_cast<const cd*>(dataChannels), reinterpret_cast<const cd*>(h0), psd,
                       data_length, nodes, K, df);
    hipLaunchKernelGGL(k_prep_reduce, dim3(ceil_div(24L * K, 256)), dim3(256), 0, st, dataConstants, partial, item, K);
    hipLaunchKernelGGL(k_prep_totals, dim3(1), dim3(192), 0, st, totals, partial, item, K);
    return hipGetLastError() == hipSuccess ? OK : ERR_LAUNCH;
}
