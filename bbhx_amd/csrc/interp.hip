// Dense-grid template build from the sparse spline coefficients
// (the reference's InterpTDI / TDI kernel, WaveformBuild.cu:41-189).
//
// The reference launches one kernel per binary on its own stream. Here all
// binaries go in ONE launch: each binary's band [start, start + n_b) is cut
// into tiles of TILE dense bins; block -> (binary, tile) by a binary search
// over the per-binary tile offsets (ragged CSR). One lane per dense bin sums
// the modes:
//   h_ch(f) = sum_m conj(T_ch,m(f) A_m(f) e^{i phi_m(f)})  if t_start <= t_m(f) <= t_end
// with every quantity a cubic in (f - f_k) on the segment k = inds[m][i].
// A first pass repacks the 36 spline coefficients and the knot of each
// (binary, mode, segment), spread over nine arrays each in the reference's
// layout, into one 320-B record (library workspace): the lanes of a wavefront
// mostly share a segment, and one base address lets the wavefront read it with
// a few wide scalar loads instead of 37 separately addressed ones.

#include <vector>

#include "bbhx_hip.h"
#include "common.h"
#include "interp_device.h"
#include "workspace.h"

using namespace bbhx;

namespace {

// Segment records are built for a chunk of binaries at a time, at most this
// many bytes (at least one binary), so that the workspace stays bounded for
// batches whose spline arrays already fill much of the device.
constexpr size_t INTERP_REC_CAP = size_t(256) << 20;

// Repack the spline arrays of segments [s0, s0 + ns) into one InterpSeg each
// (R[s - s0]): lane t -> (part p, segment s) with s fastest, so the reads of
// each spline array are unit-stride; part 9 is the knot.
__global__ __launch_bounds__(256) void k_interp_pack(InterpSeg* __restrict__ R, const double* __restrict__ freqs,
                                                     const double* __restrict__ y, const double* __restrict__ c1,
                                                     const double* __restrict__ c2, const double* __restrict__ c3,
                                                     long nper, long s0, long ns)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 10 * ns) return;
    const int p = (int)(t / ns);
    const long s = t - p * ns;
    if (p == 9) {
        R[s].x = freqs[s0 + s];
        return;
    }
    const long q = p * nper + s0 + s;
    R[s].c[p][0] = y[q];
    R[s].c[p][1] = c1[q];
    R[s].c[p][2] = c2[q];
    R[s].c[p][3] = c3[q];
}

// Tiles [tile0, tile0 + gridDim.x) of binaries [b0, b0 + nb), whose segment
// records start at R.
__global__ __launch_bounds__(INTERP_TILE) void k_interp_tdi(const InterpBin* __restrict__ bins, int b0, int nb,
                                                            int tile0, const double* __restrict__ dataFreqs,
                                                            const InterpSeg* __restrict__ R, int length,
                                                            int numModes)
{
    // e^{i phi} table in LDS: its lookups are per-lane gathers
    __shared__ double2 s_tab[129];
    fill_etab(s_tab, (int)threadIdx.x, INTERP_TILE);
    __syncthreads();
    const int tile = tile0 + (int)blockIdx.x;
    const int bl = find_bin(bins + b0, nb, tile);
    const InterpBin& d = bins[b0 + bl];
    const int i = (tile - d.tile_begin) * INTERP_TILE + threadIdx.x;
    if (i >= d.len) return;
    const double f = dataFreqs[d.start + i];
    cd h[3];
    interp_point(h, f, i, d, bl, R, length, numModes, s_tab);
    cd* out = reinterpret_cast<cd*>(d.tmpl);
    out[0 * (long)d.len + i] = h[0];
    out[1 * (long)d.len + i] = h[1];
    out[2 * (long)d.len + i] = h[2];
}

}  // namespace

namespace bbhx {
// Uploads a host table of per-binary descriptors to the table workspace.
// Synchronises the stream so the host vector may be released on return.
InterpBin* upload_table(const std::vector<InterpBin>& h, hipStream_t stream)
{
    InterpBin* d = static_cast<InterpBin*>(workspace_get(WS_TABLE, sizeof(InterpBin) * (h.size() + 1), stream));
    if (!d) return nullptr;
    if (hipMemcpyAsync(d, h.data(), sizeof(InterpBin) * h.size(), hipMemcpyHostToDevice, stream) != hipSuccess)
        return nullptr;
    if (hipStreamSynchronize(stream) != hipSuccess) return nullptr;
    return d;
}

// Host table for ragged bands cut in tiles of 'tile' dense bins; returns the
// number of tiles, or -1 when a band is invalid.
long build_table(std::vector<InterpBin>& h, const int64_t* tmpl_ptrs, const int64_t* inds_ptrs,
                 const int* inds_start, const int* ind_lengths, const double* t_start, const double* t_end,
                 int numBinAll, int tile)
{
    h.resize(numBinAll);
    long tiles = 0;
    for (int b = 0; b < numBinAll; b++) {
        if (ind_lengths[b] < 0 || inds_start[b] < 0) return -1;
        h[b].tmpl = reinterpret_cast<double*>(tmpl_ptrs ? tmpl_ptrs[b] : 0);
        h[b].inds = reinterpret_cast<const int*>(inds_ptrs ? inds_ptrs[b] : 0);
        h[b].start = inds_start[b];
        h[b].len = ind_lengths[b];
        h[b].t_start = t_start ? t_start[b] : 0.0;
        h[b].t_end = t_end ? t_end[b] : 0.0;
        h[b].tile_begin = (int)tiles;
        h[b].pad = 0;
        tiles += (ind_lengths[b] + tile - 1) / tile;
        if (tiles > 2147483647L) return -1;
    }
    return tiles;
}
}  // namespace bbhx

extern "C" int bbhx_hip_InterpTDI(const int64_t* templateChannels_ptrs, const double* dataFreqs, const double* freqs,
                                  const double* propArrays, const double* c1, const double* c2, const double* c3,
                                  const double* t_start, const double* t_end, int length, int data_length,
                                  int numBinAll, int numModes, const int64_t* inds_ptrs, const int* inds_start,
                                  const int* ind_lengths, void* stream)
{
    if (length < 2 || data_length < 0 || numBinAll < 0 || numModes < 1 || numModes > MAX_MODES) return ERR_ARGS;
    if (numBinAll == 0) return OK;
    if (!templateChannels_ptrs || !dataFreqs || !freqs || !propArrays || !c1 || !c2 || !c3 || !t_start || !t_end ||
        !inds_ptrs || !inds_start || !ind_lengths)
        return ERR_ARGS;
    for (int b = 0; b < numBinAll; b++)
        if ((long)inds_start[b] + ind_lengths[b] > data_length) return ERR_ARGS;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    std::vector<InterpBin> h;
    long tiles = build_table(h, templateChannels_ptrs, inds_ptrs, inds_start, ind_lengths, t_start, t_end, numBinAll,
                             INTERP_TILE);
    if (tiles < 0) return ERR_ARGS;
    if (tiles == 0) return OK;
    InterpBin* bins = upload_table(h, st);
    if (!bins) return ERR_ALLOC;
    const long nper = (long)numBinAll * numModes * length;
    const long per_bin = (long)numModes * length;  // segment records per binary
    long chunk = (long)(INTERP_REC_CAP / (sizeof(InterpSeg) * (size_t)per_bin));
    chunk = chunk < 1 ? 1 : (chunk > numBinAll ? numBinAll : chunk);
    InterpSeg* R = static_cast<InterpSeg*>(workspace_get(WS_INTERP, sizeof(InterpSeg) * (size_t)(per_bin * chunk), st));
    if (!R) return ERR_ALLOC;
    for (long b0 = 0; b0 < numBinAll; b0 += chunk) {
        const long b1 = b0 + chunk < numBinAll ? b0 + chunk : numBinAll;
        const int t0 = h[b0].tile_begin;
        const int t1 = b1 < numBinAll ? h[b1].tile_begin : (int)tiles;
        if (t1 == t0) continue;  // empty bands only
        const long ns = (b1 - b0) * per_bin;
        hipLaunchKernelGGL(k_interp_pack, dim3((unsigned)ceil_div(10 * ns, 256)), dim3(256), 0, st, R, freqs,
                           propArrays, c1, c2, c3, nper, b0 * per_bin, ns);
        hipLaunchKernelGGL(k_interp_tdi, dim3((unsigned)(t1 - t0)), dim3(INTERP_TILE), 0, st, bins, (int)b0,
                           (int)(b1 - b0), t0, dataFreqs, R, length, numModes);
        if (hipGetLastError() != hipSuccess) return ERR_LAUNCH;
    }
    return OK;
}
