// Sparse-grid mode sum and the inner products (the reference's
// WaveformBuild.cu:194-296 fill_waveform / direct_sum and Likelihood.cu
// hdynLikelihood, noiseweight_template, direct_like).
//
//  k_direct_sum   one lane per (binary, frequency): sum over modes
//  k_hdyn         one lane per (binary, node chunk): heterodyned sums; the
//                 (K, 3, B) residual layout is binary-fastest, so lanes over
//                 binaries read it coalesced while the constants are uniform
//  k_hdyn_reduce  sums the node chunks in a fixed order (deterministic)
//  k_direct_like  ragged tiles of one binary band; noise-weights the template
//                 in place (as the reference does) and reduces conj(d) h and
//                 |h|^2 over the tile; k_direct_like_reduce (one workgroup per
//                 binary) adds the tiles of each binary in a fixed order. No
//                 atomics anywhere.

#include <vector>

#include "bbhx_hip.h"
#include "common.h"
#include "interp_device.h"
#include "workspace.h"

using namespace bbhx;

namespace {

constexpr int SUM_THREADS = 256;

__global__ __launch_bounds__(SUM_THREADS) void k_direct_sum(cd* __restrict__ out, const double* __restrict__ buf,
                                                            int numBinAll, int L, int numModes,
                                                            const double* __restrict__ t_start,
                                                            const double* __restrict__ t_end, int b0)
{
    const int i = blockIdx.x * SUM_THREADS + threadIdx.x;
    const int b = b0 + (int)blockIdx.y;
    if (i >= L) return;
    const long nper = (long)numBinAll * numModes * L;
    const double ts = t_start[b], te = t_end[b];
    cd h[3] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
    for (int m = 0; m < numModes; m++) {
        const long idx = ((long)b * numModes + m) * L + i;
        const double tf = buf[2 * nper + idx];
        if ((tf >= ts) && ((tf <= te) || (te <= 0.0))) {
            const cd ap = buf[idx] * cexpi_red(buf[nper + idx]);
            h[0] += conj(mk(buf[3 * nper + idx], buf[4 * nper + idx]) * ap);
            h[1] += conj(mk(buf[5 * nper + idx], buf[6 * nper + idx]) * ap);
            h[2] += conj(mk(buf[7 * nper + idx], buf[8 * nper + idx]) * ap);
        }
    }
    for (int c = 0; c < 3; c++) out[((long)b * 3 + c) * L + i] = h[c];
}

// Heterodyned sums over segments j in [j0, j1) of one channel range, per binary.
// Segment j joins nodes j-1 and j and uses the constants at index j.
__global__ __launch_bounds__(64) void k_hdyn(cd* __restrict__ partial, const cd* __restrict__ r,
                                             const cd* __restrict__ dc, const double* __restrict__ f, int numBinAll,
                                             int K, int nch, int jend, int chunk)
{
    const int b = blockIdx.x * 64 + threadIdx.x;
    const int g = blockIdx.y;  // node chunk
    if (b >= numBinAll) return;
    const int j0 = 1 + g * chunk;
    const int j1 = (j0 + chunk < jend) ? j0 + chunk : jend;
    cd l1 = mk(0.0, 0.0), l2 = mk(0.0, 0.0);
    for (int ch = 0; ch < nch; ch++) {
        if (j0 >= j1) break;
        cd prev = r[((long)(j0 - 1) * nch + ch) * numBinAll + b];
        double fprev = f[j0 - 1];
        for (int j = j0; j < j1; j++) {
            const double fj = f[j];
            const cd cur = r[((long)j * nch + ch) * numBinAll + b];
            if (fprev != 0.0) {
                const cd A0 = dc[(0 * nch + ch) * K + j];
                const cd A1 = dc[(1 * nch + ch) * K + j];
                const cd B0 = dc[(2 * nch + ch) * K + j];
                const cd B1 = dc[(3 * nch + ch) * K + j];
                const double inv = 1.0 / (fj - fprev);
                const cd r1 = mk((cur.re - prev.re) * inv, (cur.im - prev.im) * inv);
                const double fm = (fj + fprev) / 2.0;
                const cd r0 = cur - r1 * (fj - fm);
                const cd r1c = conj(r1);
                l1 += A0 * conj(r0) + A1 * r1c;
                const double mag2 = r0.re * r0.re + r0.im * r0.im;
                l2 += B0 * mag2 + 2. * B1 * (r0 * r1c).re;
            }
            prev = cur;
            fprev = fj;
        }
    }
    partial[(long)g * 2 * numBinAll + b] = l1;
    partial[(long)g * 2 * numBinAll + numBinAll + b] = l2;
}

__global__ void k_hdyn_reduce(cd* __restrict__ o1, cd* __restrict__ o2, const cd* __restrict__ partial,
                              int numBinAll, int nchunk)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= numBinAll) return;
    cd s1 = mk(0.0, 0.0), s2 = mk(0.0, 0.0);
    for (int g = 0; g < nchunk; g++) {
        s1 += partial[(long)g * 2 * numBinAll + b];
        s2 += partial[(long)g * 2 * numBinAll + numBinAll + b];
    }
    o1[b] = s1;
    o2[b] = s2;
}

constexpr int DL_THREADS = 256;
constexpr int DL_PER_THREAD = 8;
constexpr int DL_TILE = DL_THREADS * DL_PER_THREAD;

__device__ inline double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide sum of 3 doubles; result valid in thread 0.
__device__ inline void block_sum3(double& a, double& b, double& c)
{
    __shared__ double s[3][DL_THREADS / 64];
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        s[0][w] = a;
        s[1][w] = b;
        s[2][w] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = b = c = 0.0;
        for (int k = 0; k < DL_THREADS / 64; k++) {
            a += s[0][k];
            b += s[1][k];
            c += s[2][k];
        }
    }
}

// bins[b].tile_begin counts DL_TILE tiles here.
__global__ __launch_bounds__(DL_THREADS) void k_direct_like(double* __restrict__ partial,
                                                            const InterpBin* __restrict__ bins, int numBinAll,
                                                            const cd* __restrict__ data,
                                                            const double* __restrict__ w, int N)
{
    const int b = find_bin(bins, numBinAll, (int)blockIdx.x);
    const InterpBin& d = bins[b];
    cd* t = reinterpret_cast<cd*>(d.tmpl);
    const int base = ((int)blockIdx.x - d.tile_begin) * DL_TILE;
    double dh_re = 0.0, dh_im = 0.0, hh = 0.0;
    for (int k = 0; k < DL_PER_THREAD; k++) {
        const int i = base + k * DL_THREADS + threadIdx.x;
        if (i >= d.len) break;
        for (int c = 0; c < 3; c++) {
            const long di = (long)c * N + d.start + i;
            cd h = t[(long)c * d.len + i] * w[di];
            t[(long)c * d.len + i] = h;
            const cd dd = data[di];
            dh_re += dd.re * h.re + dd.im * h.im;
            dh_im += dd.re * h.im - dd.im * h.re;
            hh += h.re * h.re + h.im * h.im;
        }
    }
    block_sum3(dh_re, dh_im, hh);
    if (threadIdx.x == 0) {
        partial[3 * (long)blockIdx.x + 0] = dh_re;
        partial[3 * (long)blockIdx.x + 1] = dh_im;
        partial[3 * (long)blockIdx.x + 2] = hh;
    }
}

// One workgroup per binary: its threads take the binary's tiles in a fixed
// stride and the workgroup adds the per-thread sums in a fixed tree (a band of
// 12.6 M bins is ~6000 tiles, too many for one thread's dependent loads).
__global__ __launch_bounds__(DL_THREADS) void k_direct_like_reduce(cd* __restrict__ d_h, cd* __restrict__ h_h,
                                                                   const double* __restrict__ partial,
                                                                   const InterpBin* __restrict__ bins, int numBinAll,
                                                                   int total_tiles)
{
    const int b = blockIdx.x;
    const int t0 = bins[b].tile_begin;
    const int t1 = (b + 1 < numBinAll) ? bins[b + 1].tile_begin : total_tiles;
    double a = 0.0, c = 0.0, e = 0.0;
    for (int t = t0 + (int)threadIdx.x; t < t1; t += DL_THREADS) {
        a += partial[3 * (long)t + 0];
        c += partial[3 * (long)t + 1];
        e += partial[3 * (long)t + 2];
    }
    block_sum3(a, c, e);
    if (threadIdx.x == 0) {
        d_h[b] = d_h[b] + mk(4.0 * a, 4.0 * c);
        h_h[b] = h_h[b] + mk(4.0 * e, 0.0);
    }
}

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace

extern "C" {

int bbhx_hip_direct_sum(double* templateChannels, const double* bbh_buffer, int numBinAll, int data_length,
                        int nChannels, int numModes, const double* t_start, const double* t_end, void* stream)
{
    if (numBinAll < 0 || data_length < 1 || nChannels != 3 || numModes < 1 || numModes > MAX_MODES) return ERR_ARGS;
    if (numBinAll == 0) return OK;
    if (!templateChannels || !bbh_buffer || !t_start || !t_end) return ERR_ARGS;
    for (int b0 = 0; b0 < numBinAll; b0 += 65535) {
        int nb = numBinAll - b0 < 65535 ? numBinAll - b0 : 65535;
        hipLaunchKernelGGL(k_direct_sum, dim3(ceil_div(data_length, SUM_THREADS), nb), dim3(SUM_THREADS), 0,
                           S(stream), reinterpret_cast<cd*>(templateChannels), bbh_buffer, numBinAll, data_length,
                           numModes, t_start, t_end, b0);
    }
    return hipGetLastError() == hipSuccess ? OK : ERR_LAUNCH;
}

int bbhx_hip_hdyn(double* likeOut1, double* likeOut2, const double* templateChannels, const double* dataConstants,
                  const double* dataFreqs, int numBinAll, int data_length, int nChannels, int include_last_segment,
                  void* stream)
{
    if (numBinAll < 0 || data_length < 2 || nChannels < 1 || nChannels > 3) return ERR_ARGS;
    if (numBinAll == 0) return OK;
    if (!likeOut1 || !likeOut2 || !templateChannels || !dataConstants || !dataFreqs) return ERR_ARGS;
    const int jend = include_last_segment ? data_length : data_length - 1;
    // enough node chunks to give the device ~4 waves per CU
    int waves = ceil_div(numBinAll, 64);
    int nchunk = waves >= 1024 ? 1 : ceil_div(1024, waves);
    if (nchunk > 64) nchunk = 64;
    int chunk = ceil_div(jend - 1 > 0 ? jend - 1 : 1, nchunk);
    nchunk = ceil_div(jend - 1 > 0 ? jend - 1 : 1, chunk);
    cd* partial = static_cast<cd*>(workspace_get(WS_PARTIAL, sizeof(cd) * 2 * (size_t)numBinAll * nchunk, S(stream)));
    if (!partial) return ERR_ALLOC;
    hipLaunchKernelGGL(k_hdyn, dim3(waves, nchunk), dim3(64), 0, S(stream), partial,
                       reinterpret_cast<const cd*>(templateChannels), reinterpret_cast<const cd*>(dataConstants),
                       dataFreqs, numBinAll, data_length, nChannels, jend, chunk);
    if (hipGetLastError() != hipSuccess) return ERR_LAUNCH;
    hipLaunchKernelGGL(k_hdyn_reduce, dim3(ceil_div(numBinAll, 256)), dim3(256), 0, S(stream),
                       reinterpret_cast<cd*>(likeOut1), reinterpret_cast<cd*>(likeOut2), partial, numBinAll, nchunk);
    return hipGetLastError() == hipSuccess ? OK : ERR_LAUNCH;
}

int bbhx_hip_direct_like(double* d_h, double* h_h, const double* dataChannels, const double* noise_weight_times_df,
                         const int64_t* templateChannels_ptrs, const int* inds_start, const int* ind_lengths,
                         int data_stream_length, int numBinAll, void* stream)
{
    if (numBinAll < 0 || data_stream_length < 1) return ERR_ARGS;
    if (numBinAll == 0) return OK;
    if (!d_h || !h_h || !dataChannels || !noise_weight_times_df || !templateChannels_ptrs || !inds_start ||
        !ind_lengths)
        return ERR_ARGS;
    for (int b = 0; b < numBinAll; b++)
        if (inds_start[b] < 0 || ind_lengths[b] < 0 || (long)inds_start[b] + ind_lengths[b] > data_stream_length)
            return ERR_ARGS;
    hipStream_t st = S(stream);
    std::vector<InterpBin> h;
    long tiles = build_table(h, templateChannels_ptrs, nullptr, inds_start, ind_lengths, nullptr, nullptr, numBinAll,
                             DL_TILE);
    if (tiles < 0) return ERR_ARGS;
    if (tiles == 0) return OK;
    InterpBin* bins = upload_table(h, st);
    if (!bins) return ERR_ALLOC;
    double* partial = static_cast<double*>(workspace_get(WS_PARTIAL, sizeof(double) * 3 * (size_t)tiles, st));
    if (!partial) return ERR_ALLOC;
    hipLaunchKernelGGL(k_direct_like, dim3((unsigned)tiles), dim3(DL_THREADS), 0, st, partial, bins, numBinAll,
                       reinterpret_cast<const cd*>(dataChannels), noise_weight_times_df, data_stream_length);
    if (hipGetLastError() != hipSuccess) return ERR_LAUNCH;
    hipLaunchKernelGGL(k_direct_like_reduce, dim3(numBinAll), dim3(DL_THREADS), 0, st,
                       reinterpret_cast<cd*>(d_h), reinterpret_cast<cd*>(h_h), partial, bins, numBinAll, (int)tiles);
    return hipGetLastError() == hipSuccess ? OK : ERR_LAUNCH;
}

}  // extern "C"
