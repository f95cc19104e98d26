// Fused direct log-likelihood of a ragged template batch on gfx950 (no
// counterpart entry point in the reference; it computes what
// Likelihood.get_ll does, likelihood.py:140-225, for templates whose sparse
// frequency grids have different lengths, in one launch sequence with no host
// synchronisation, so a caller can capture it in a hipGraph).
//
// Per binary b with sparse length L_b (layout: points of all binaries
// concatenated, offset off_b; within a binary [i][mode] for the grid and
// [i][mode] segment records: the knots x_i, x_{i+1} that bound segment i, then
// per quantity [y, c1, c2, c3], so that one segment of one quantity is 32
// contiguous bytes and one record read both checks and evaluates the segment):
//   k_dir_setup   one lane per (binary, mode): coefficient records
//   k_dir_grid    one wavefront per (binary, mode): default grid
//                 Mf in [mf_min, mf_max] * m / 2, t(f) there, the crossings of
//                 t = 0 and t = Tobs, the grid refitted to that window
//                 (phenomhm.py:208-224, :509-614), then amplitude, phase, t(f)
//                 and the TDI response on the refitted grid (phenomhm.py:487-507,
//                 Response.cu), written compact: nine values per (knot, mode)
//   k_dir_spline  one wavefront per binary, one lane per (mode, quantity):
//                 not-a-knot spline coefficients (Interpolate.cu:15-359) from
//                 the compact values (forward sweep into the scratch Fs, then
//                 each record's (y, c1, c2, c3) written together, once)
//   k_dir_band    dense band [st, et) of each binary (waveformbuild.py:139-156)
//                 and the exclusive prefix of its chunks of dense bins (chunks
//                 aligned to absolute multiples of the chunk size)
//   k_dir_cnt_zero, k_dir_cnt, k_dir_cnt_scan, k_dir_perm
//                 chunk-major execution order of the work items (counting sort)
//   k_dir_like    persistent: workgroups loop over the (binary, chunk) work
//                 items in that order; each thread evaluates dense bins (spline segment per
//                 mode, WaveformBuild.cu:80-143, window and mode sum :48-60)
//                 and accumulates the noise-weighted inner products
//                 (Likelihood.cu:357-471); one partial sum per work item
//   k_dir_reduce  one wavefront per binary: its partials in a fixed order (deterministic)

#include <algorithm>
#include <type_traits>
#include <vector>

#include "bbhx_hip.h"
#include "common.h"
#include "fused_device.h"
#include "grad_device.h"

using namespace bbhx;
using namespace bbhx::fused;

namespace bbhx {
bool response_orbit_tab(const bbhx_response* r, resp::OrbitTab& o);
}

struct bbhx_direct_plan {
    int device = -1;
    int B = 0, M = 0, N = 0, Lmax = 0;
    int min_chunk = 0;   // smallest chunk the partial-sum buffer is sized for
    long S = 0;          // total sparse points
    long max_items = 0;  // bound on (binary, chunk) work items
    // device buffers
    long* off = nullptr;   // (B,) first point of each binary
    int* len = nullptr;    // (B,)
    FusedRec* recs = nullptr;
    double* X = nullptr;   // (S, M)
    double* Cf = nullptr;  // (S, M, RS): segment records (bounds, value and cubic coefficients)
    double* Yc = nullptr;  // (S, M, NQ): the grid pass's nine values per knot, compact
    double* Fs = nullptr;  // (S, M, NQ + 1): the spline's forward sweep, d' of each quantity, then c'
    double* grid_lo = nullptr;    // (B, M) log2 of the first refitted grid point
    double* grid_istep = nullptr; // (B, M) 1 / log2 step (0 for a degenerate grid)
    int* band = nullptr;          // (B, 3): st, et, chunk_begin
    int* total = nullptr;         // (1,) number of work items
    double* partial = nullptr;    // (max_items, 4)
    int* perm = nullptr;          // (max_items,) execution order of the work items
    int* cnt = nullptr;           // (max_chunks + 1,) items per absolute chunk, then offsets
    int* next = nullptr;          // (1,) next entry of perm for the dense-band kernel
    // gradient buffers (bbhx_hip_direct_plan_enable_grad)
    double* Yadj = nullptr;              // (S, M, 9, 4) adjoint records
    bbhx::gradk::GradRec* grecs = nullptr;  // (B, M) records with tangents
    double* gpart = nullptr;             // (B, M, 11) per-mode gradients
    // FP32-bulk form (bbhx_hip_direct_plan_set_precision)
    int precision = 0;                   // 0: FP64, 1: FP32 bulk, FP64 phase and t(f)
    void* CfF = nullptr;                 // (S, M) 192-B segment records of that form
    int* loud = nullptr;                 // (max_items,) work items of the binaries evaluated in FP64
    int* quiet = nullptr;                // (max_items,) work items of the binaries evaluated in FP32
    int* loud_cnt = nullptr;             // (4,) loud count, its pass's next-item counter; the same for quiet
    int* route = nullptr;                // (B,) 1: the h_h estimate sends the binary to the FP64 pass
    int* split_cnt = nullptr;            // (max_items / 256 + 1,) loud entries per block of the order, scanned
    int max_chunks = 0;           // absolute chunks of min_chunk bins
};

namespace {

constexpr int NQ = 9;


// one spline segment of one quantity: value at the knot and c1, c2, c3
struct alignas(32) Cubic {
    double y, c1, c2, c3;
};

// segment i of one (binary, mode): it covers x_lo <= f < x_hi (x_hi = x_lo for
// the last knot, which starts no segment)
struct alignas(32) SegRec {
    double x_lo, x_hi, pad0, pad1;
    Cubic q[NQ];
};
constexpr int RS = sizeof(SegRec) / sizeof(double);  // 40 doubles
constexpr int RQ = 4;                                 // offset of q[0] in doubles
static_assert(sizeof(SegRec) == 320, "segment record layout");

// Segment record of the FP32-bulk form: bounds and the phase and t(f) cubics
// in FP64 (phases reach 1e8 rad, and t(f) decides the observation window
// exactly as in the FP64 form), amplitude and the six transfer-function parts
// (T1..T3 re, im) in FP32.
struct alignas(64) SegRecF {
    double x_lo, x_hi;
    double ph[4];    // y, c1, c2, c3
    double tf[4];
    float q[7][4];   // A, T1 re, T1 im, T2 re, T2 im, T3 re, T3 im: y, c1, c2, c3
};
static_assert(sizeof(SegRecF) == 192, "FP32-bulk segment record layout");

// single-precision complex (the mode sum of the FP32-bulk form)
struct cf {
    float re, im;
};

__device__ inline cd to_cd(cd a) { return a; }
__device__ inline cd to_cd(cf a) { return mk((double)a.re, (double)a.im); }
__device__ inline void zero(cd& a) { a = mk(0.0, 0.0); }
__device__ inline void zero(cf& a) { a = cf{0.0f, 0.0f}; }

__device__ inline const Cubic& cubic_of(const SegRec& r, int k) { return r.q[k]; }
__device__ inline const Cubic& cubic_of(const Cubic (&c)[NQ], int k) { return c[k]; }

__global__ __launch_bounds__(64) void k_dir_setup(FusedRec* __restrict__ recs, const double* __restrict__ params, int B, ModeList modes,
                            int M, const double* __restrict__ qnm, int table_len, double dspin)
{
    setup_split(recs, params, B, modes, M, qnm, table_len, dspin);
}

// numpy.logspace point i of num between 10^lo and 10^hi (start + i step, the
// last point exactly 10^hi)
__device__ inline double logspace_pt(double lo, double step, double hi, int i, int num)
{
    return pow(10.0, i == num - 1 ? hi : lo + i * step);
}

__device__ inline int wsum_i(int v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline int wmax_i(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ inline double wmin_d(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ inline double wmax_d(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// One 64-lane wavefront per (binary, mode). LDS: t(f) of the default grid (L_b doubles).
__global__ __launch_bounds__(64) void k_dir_grid(double* __restrict__ X, double* __restrict__ Yc,
                                                 double* __restrict__ grid_lo, double* __restrict__ grid_istep,
                                                 int* __restrict__ flags, const FusedRec* __restrict__ recs,
                                                 const long* __restrict__ off, const int* __restrict__ len,
                                                 const double* __restrict__ params,
                                                 const double* __restrict__ t_start,
                                                 const double* __restrict__ t_end, int B, int M, ModeList modes,
                                                 resp::OrbitTab orb, int TDItag, int rescaled, int tdi2,
                                                 double initial_t_val, double mf_min, double mf_max)
{
    extern __shared__ double tfs[];
    const int bm = blockIdx.x;
    const int b = bm / M, m = bm % M;
    const int lane = threadIdx.x;
    const int L = len[b];
    const long base = off[b];
    const FusedRec& r = recs[bm];
    const double t_ref = params[P_TREF * B + b];
    const double Tobs = t_end[b] - t_start[b];
    const double Msec = (params[P_M1 * B + b] + params[P_M2 * B + b]) * MTSUN_SI;
    const double mm = modes.mm[m];

    // default grid (phenomhm.py:208-224) and t(f) on it; the frequencies go to X for now
    const double lo = log10(mf_min), hi = log10(mf_max), step = (hi - lo) / (L - 1);
    for (int i = lane; i < L; i += 64) {
        const double f = logspace_pt(lo, step, hi, i, L) / Msec * mm / 2;
        double amp, phase, tf;
        phm::phm_eval(r.w, f, amp, phase, tf);
        tfs[i] = tf + t_ref;
        X[(base + i) * M + m] = f;
    }
    __syncthreads();

    // window search (phenomhm.py:520-614): the single segment where t crosses 0 / Tobs
    double tmin = 1e300, tmax = -1e300;
    int n0 = 0, nT = 0, s0 = -1, sT = -1;
    for (int i = lane; i < L; i += 64) {
        const double t = tfs[i];
        tmin = fmin(tmin, t);
        tmax = fmax(tmax, t);
        if (i + 1 < L) {
            const double t1 = tfs[i + 1];
            if ((0.0 < t1) && (0.0 > t)) {
                n0++;
                s0 = i;
            }
            if ((Tobs < t1) && (Tobs > t)) {
                nT++;
                sT = i;
            }
        }
    }
    tmin = wmin_d(tmin);
    tmax = wmax_d(tmax);
    n0 = wsum_i(n0);
    nT = wsum_i(nT);
    s0 = wmax_i(s0);
    sT = wmax_i(sT);
    const double* xf = X + base * M + m;
    double f_min = xf[0], f_max = xf[(long)(L - 1) * M];
    int flag = 0;
    if (tmin < 0.0) {
        if (n0 != 1) flag |= 1;
        else {
            const double y1 = xf[(long)s0 * M], y2 = xf[(long)(s0 + 1) * M];
            const double x1 = tfs[s0], x2 = tfs[s0 + 1];
            f_min = (y2 - y1) / (x2 - x1) * (0.0 - x1) + y1;
        }
    }
    if (tmax > Tobs) {
        if (nT != 1) flag |= 2;
        else {
            const double y1 = xf[(long)sT * M], y2 = xf[(long)(sT + 1) * M];
            const double x1 = tfs[sT], x2 = tfs[sT + 1];
            f_max = (y2 - y1) / (x2 - x1) * (Tobs - x1) + y1;
        }
    }
    if (lane == 0 && flag) atomicOr(&flags[b], flag);
    __syncthreads();  // every lane has read the default grid before it is overwritten

    // refitted grid and the nine quantities on it
    const double lo2 = log10(f_min), hi2 = log10(f_max), step2 = (hi2 - lo2) / (L - 1);
    if (lane == 0) {
        // log2 units: the dense kernel estimates the segment from an FP32 log2
        grid_lo[bm] = lo2 * LOG2_10;
        grid_istep[bm] = step2 > 0.0 ? 1.0 / (step2 * LOG2_10) : 0.0;
    }
    const double dt_shift = t_ref - initial_t_val;
    for (int i = lane; i < L; i += 64) {
        const double f = logspace_pt(lo2, step2, hi2, i, L);
        double amp, phase, tf;
        phm::phm_eval(r.w, f, amp, phase, tf);
        phase = phase + ((f * dt_shift) * 2) * PI;
        tf = tf + t_ref;
        cd T[3];
        double dphi;
        resp::resp_eval(r.s, orb, f, tf, TDItag, rescaled, tdi2, T, dphi);
        phase = phase + dphi;
        const long p = (base + i) * M + m;
        X[p] = f;
        double* y = Yc + p * NQ;
        y[0] = amp;
        y[1] = phase;
        y[2] = tf;
        y[3] = T[0].re;
        y[4] = T[0].im;
        y[5] = T[1].re;
        y[6] = T[1].im;
        y[7] = T[2].re;
        y[8] = T[2].im;
    }
}

constexpr int DSPL_BLOCK = 8;  // rows loaded ahead of the recurrence in each sweep of k_dir_spline

// Not-a-knot spline of one binary: lane s = mode * 9 + quantity runs the
// Thomas algorithm over the binary's L knots (the system of spline.hip);
// consecutive lanes read consecutive quantities of the same knot.
__global__ __launch_bounds__(64) void k_dir_spline(const double* __restrict__ X, const double* __restrict__ Yc,
                                                   double* __restrict__ Cf, double* __restrict__ Fs,
                                                   SegRecF* __restrict__ CfF, const long* __restrict__ off,
                                                   const int* __restrict__ len, int M)
{
    const int b = blockIdx.x;
    const int s = threadIdx.x;
    if (s >= M * NQ) return;
    const int m = s / NQ;
    const int n = len[b];
    const long base = off[b];
    // element i of this system
    auto xi = [&](int i) { return X[(base + i) * M + m]; };
    const int q = s - m * NQ;
    auto at = [&](int i) { return ((base + i) * M + m) * RS + RQ + q * 4; };
    auto yi = [&](int i) { return Yc[((base + i) * M + m) * NQ + q]; };
    // the forward sweep's d' of this quantity and the mode's c' (the same for
    // its nine quantities: it depends on the knots alone), NQ + 1 per (knot, mode)
    auto fs = [&](int i) { return ((base + i) * M + m) * (NQ + 1); };
    // the FP32-bulk record of point i (CfF != nullptr): this lane's quantity
    auto put_f = [&](int i, double x_lo, double x_hi, double c0, double c1, double c2, double c3) {
        SegRecF& rf = CfF[(base + i) * M + m];
        if (q == 1) {
            rf.ph[0] = c0;
            rf.ph[1] = c1;
            rf.ph[2] = c2;
            rf.ph[3] = c3;
        } else if (q == 2) {
            rf.tf[0] = c0;
            rf.tf[1] = c1;
            rf.tf[2] = c2;
            rf.tf[3] = c3;
        } else {
            const int k = q == 0 ? 0 : q - 2;
            rf.q[k][0] = (float)c0;
            rf.q[k][1] = (float)c1;
            rf.q[k][2] = (float)c2;
            rf.q[k][3] = (float)c3;
        }
        if (q == 0) {
            rf.x_lo = x_lo;
            rf.x_hi = x_hi;
        }
    };

    // The forward sweep writes only d' and c' (80 B per (knot, mode), the
    // wavefront's lanes consecutive) to the scratch Fs; the backward sweep
    // writes every record once, in whole 32-B groups.
    auto put_fwd = [&](int i, double d, double c) {
        Fs[fs(i) + q] = d;
        if (q == 0) Fs[fs(i) + NQ] = c;
    };
    double x0 = xi(0), x1 = xi(1), x2 = xi(2);
    double y0 = yi(0), y1 = yi(1), y2 = yi(2);
    double cprev, dprev;
    {
        double dx1 = x1 - x0, dx2 = x2 - x1, dd = x2 - x0;
        double s1 = (y1 - y0) / dx1, s2 = (y2 - y1) / dx2;
        double bb = ((dx1 + 2 * dd) * dx2 * s1 + dx1 * dx1 * s2) / dd;
        cprev = dd / dx2;
        dprev = bb / dx2;
        put_fwd(0, dprev, cprev);
    }
    // each sweep loads DSPL_BLOCK rows ahead of its recurrence (one memory
    // latency per block) and divides once per row
    double xa = x0, xb = x1, ya = y0, yb = y1;
    for (int i0 = 1; i0 < n - 1; i0 += DSPL_BLOCK) {
        double xv[DSPL_BLOCK], yv[DSPL_BLOCK];
#pragma unroll
        for (int u = 0; u < DSPL_BLOCK; u++) {
            const int i = i0 + u;
            xv[u] = i < n - 1 ? xi(i + 1) : 0.0;
            yv[u] = i < n - 1 ? yi(i + 1) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < DSPL_BLOCK; u++) {
            const int i = i0 + u;
            if (i >= n - 1) break;
            const double xc = xv[u], yc = yv[u];
            const double ex1 = xb - xa, ex2 = xc - xb;
            const double t1 = (yb - ya) / ex1, t2 = (yc - yb) / ex2;
            const double rhs = 3.0 * (ex2 * t1 + ex1 * t2);
            const double diag = 2 * (ex1 + ex2), lo = ex2, up = ex1;
            const double im = 1.0 / (diag - lo * cprev);
            cprev = up * im;
            dprev = (rhs - lo * dprev) * im;
            put_fwd(i, dprev, cprev);
            xa = xb;
            xb = xc;
            ya = yb;
            yb = yc;
        }
    }
    double snext;  // the last row's slope, the start of the backward sweep
    {
        const double xm3 = xi(n - 3), xm2 = xi(n - 2), xm1 = xi(n - 1);
        const double ym3 = yi(n - 3), ym2 = yi(n - 2), ym1 = yi(n - 1);
        const double ex1 = xm2 - xm3, ex2 = xm1 - xm2, ed = xm1 - xm3;
        const double t1 = (ym2 - ym3) / ex1, t2 = (ym1 - ym2) / ex2;
        const double rhs = ((ex2 * ex2 * t1 + (2 * ed + ex2) * ex1 * t2) / ed);
        const double mm = ex1 - ed * cprev;
        snext = (rhs - ed * dprev) / mm;
    }
    double xn = xi(n - 1), yn = yi(n - 1);
    *reinterpret_cast<Cubic*>(Cf + at(n - 1)) = Cubic{yn, 0.0, 0.0, 0.0};
    if (q == 0) *reinterpret_cast<Cubic*>(Cf + ((base + n - 1) * M + m) * RS) = Cubic{xn, xn, 0.0, 0.0};
    if (CfF) put_f(n - 1, xn, xn, yn, 0.0, 0.0, 0.0);
    for (int i0 = n - 2; i0 >= 0; i0 -= DSPL_BLOCK) {
      double dv[DSPL_BLOCK], cv[DSPL_BLOCK], xv[DSPL_BLOCK], yv[DSPL_BLOCK];
#pragma unroll
      for (int u = 0; u < DSPL_BLOCK; u++) {
          const int i = i0 - u;
          dv[u] = i >= 0 ? Fs[fs(i) + q] : 0.0;
          cv[u] = i >= 0 ? Fs[fs(i) + NQ] : 0.0;
          xv[u] = i >= 0 ? xi(i) : 0.0;
          yv[u] = i >= 0 ? yi(i) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < DSPL_BLOCK; u++) {
        const int i = i0 - u;
        if (i < 0) break;
        const double si = dv[u] - cv[u] * snext;
        const double x_i = xv[u], y_i = yv[u];
        const double idx = 1.0 / (xn - x_i);
        const double slope = (yn - y_i) * idx;
        const double t = (si + snext - 2 * slope) * idx;
        const double c2 = (slope - si) * idx - t, c3 = t * idx;
        *reinterpret_cast<Cubic*>(Cf + at(i)) = Cubic{y_i, si, c2, c3};
        if (CfF) put_f(i, x_i, xn, y_i, si, c2, c3);
        if (q == 0) *reinterpret_cast<Cubic*>(Cf + ((base + i) * M + m) * RS) = Cubic{x_i, xn, 0.0, 0.0};
        snext = si;
        xn = x_i;
        yn = y_i;
      }
    }
}

// number of entries of the sorted array a[0..n) that are <= v (searchsorted side='right')
__device__ inline int search_right(const double* __restrict__ a, int n, double v)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Dense band of each binary and the exclusive prefix of its chunks; one
// workgroup of 1024 threads, each owning a contiguous run of binaries. Chunks
// are aligned to absolute multiples of 'chunk' dense bins, so that the work
// items of different binaries over the same dense bins read the same data.
__global__ __launch_bounds__(1024) void k_dir_band(int* __restrict__ band, int* __restrict__ total,
                                                   const double* __restrict__ X, const long* __restrict__ off,
                                                   const int* __restrict__ len, const double* __restrict__ freqs,
                                                   int N, int B, int M, int chunk)
{
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (B + 1023) / 1024;
    const int b0 = t * per, b1 = min(B, b0 + per);
    int sum = 0;
    for (int b = b0; b < b1; b++) {
        const long base = off[b];
        const int L = len[b];
        double lo = 1e300, hi = -1e300;
        for (int m = 0; m < M; m++) {
            lo = fmin(lo, X[base * M + m]);
            hi = fmax(hi, X[(base + L - 1) * M + m]);
        }
        const int st = search_right(freqs, N, lo), et = search_right(freqs, N, hi);
        const int nch = et > st ? (et - 1) / chunk - st / chunk + 1 : 0;
        band[3 * b] = st;
        band[3 * b + 1] = et;
        band[3 * b + 2] = nch;  // count for now
        sum += nch;
    }
    part[t] = sum;
    __syncthreads();
    // inclusive scan of the 1024 partial counts (Hillis-Steele)
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = t > 0 ? part[t - 1] : 0;
    for (int b = b0; b < b1; b++) {
        const int nch = band[3 * b + 2];
        band[3 * b + 2] = run;
        run += nch;
    }
    if (t == 1023) total[0] = part[1023];
}

// Execution order of the work items, chunk-major: perm lists the items of
// absolute chunk 0, then of chunk 1, ... (a counting sort; the order inside a
// chunk is whatever the atomics give, which does not affect any result: each
// item writes its own partial, reduced per binary in item order). Concurrent
// workgroups then work on the same dense bins for different binaries, so the
// data and noise weights of a chunk come from L2 instead of being fetched again
// for every binary. cnt has nchunks + 1 entries.
__global__ void k_dir_cnt_zero(int* __restrict__ cnt, int n, int* __restrict__ next)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cnt[i] = 0;
    if (i == 0) next[0] = 0;
}

// binary owning work item 'it' (largest b with chunk_begin <= it)
__device__ inline int owner(const int* __restrict__ band, int B, int it)
{
    int lo = 0, hi = B - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (band[3 * mid + 2] <= it) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// absolute chunk of work item it
__device__ inline int item_chunk(const int* __restrict__ band, int B, int it, int chunk, int& b)
{
    b = owner(band, B, it);
    return band[3 * b] / chunk + (it - band[3 * b + 2]);
}

// one lane per work item
__global__ void k_dir_cnt(int* __restrict__ cnt, const int* __restrict__ band, const int* __restrict__ total, int B,
                          int chunk)
{
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= total[0]) return;
    int b;
    atomicAdd(&cnt[item_chunk(band, B, it, chunk, b)], 1);
}

// exclusive scan of cnt[0, n) in place; one workgroup of 1024 threads
__global__ __launch_bounds__(1024) void k_dir_cnt_scan(int* __restrict__ cnt, int n)
{
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int g0 = t * per, g1 = min(n, g0 + per);
    int sum = 0;
    for (int g = g0; g < g1; g++) sum += cnt[g];
    part[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = t > 0 ? part[t - 1] : 0;
    for (int g = g0; g < g1; g++) {
        const int c = cnt[g];
        cnt[g] = run;
        run += c;
    }
}

__global__ void k_dir_perm(int* __restrict__ perm, int* __restrict__ cnt, const int* __restrict__ band,
                           const int* __restrict__ total, int B, int chunk)
{
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= total[0]) return;
    int b;
    perm[atomicAdd(&cnt[item_chunk(band, B, it, chunk, b)], 1)] = it;
}

// h[ch] += conj(T_ch A e^{i phase}) of one mode at x = f - x_i into its segment, if
// t_start <= tf <= t_end (t_end <= 0: no upper limit) (WaveformBuild.cu:48-60, :80-143)
template <class Q>
__device__ inline void add_mode(cd h[3], const Q& q, double x, double ts, double te, const double2* tab)
{
    const double x2 = x * x, x3 = x * x2;
    auto ev = [&](int k) {
        const Cubic c = cubic_of(q, k);
        return c.y + c.c1 * x + c.c2 * x2 + c.c3 * x3;
    };
    const double tf = ev(2);
    if ((tf >= ts) && ((tf <= te) || (te <= 0.0))) {
        const cd ap = ev(0) * cexpi_tab(ev(1), tab);
        h[0] += conj(mk(ev(3), ev(4)) * ap);
        h[1] += conj(mk(ev(5), ev(6)) * ap);
        h[2] += conj(mk(ev(7), ev(8)) * ap);
    }
}

// Segment of one lane's dense frequency f in the grid of one mode, searched by
// that lane, and its contribution (lanes of a wavefront in many different
// segments). The record at the FP32 log2 estimate is read with its bounds and
// all nine cubics in one round of loads; the index then moves until
// x_lo <= f < x_hi or leaves [0, L - 2] (no contribution, WaveformBuild.cu:92-93).
__device__ inline void mode_per_lane(cd h[3], const SegRec* __restrict__ rm, int M, int L, double f, int i, double ts,
                                     double te, const double2* tab)
{
    const SegRec* r = rm + (long)i * M;
    double2 bnd = *reinterpret_cast<const double2*>(r);
    Cubic cq[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) cq[q] = r->q[q];
    // keeps the compiler from sinking these loads below the bounds check
    // (which would serialise two rounds of memory latency)
#pragma unroll
    for (int q = 0; q < NQ; q++) asm volatile("" ::"v"(cq[q].y), "v"(cq[q].c1), "v"(cq[q].c2), "v"(cq[q].c3));
    if (!(bnd.x <= f && f < bnd.y)) {
        // the estimate missed (rounding near a knot): walk to the segment holding f
        while (f < bnd.x && i > 0) bnd = *reinterpret_cast<const double2*>(rm + (long)--i * M);
        while (f >= bnd.y && i < L - 2) bnd = *reinterpret_cast<const double2*>(rm + (long)++i * M);
        if (!(bnd.x <= f && f < bnd.y)) return;
        r = rm + (long)i * M;
#pragma unroll
        for (int q = 0; q < NQ; q++) cq[q] = r->q[q];
    }
    add_mode(h, cq, f - bnd.x, ts, te, tab);
}

// e^{i phase} in FP32 from an FP64 phase: the phase in revolutions reduced to
// [-1/2, 1/2] in FP64 (for |phase| <= 1e9 rad the product and the rounding cost
// < 2e-8 rad), then the hardware single-precision cosine and sine, which take
// their argument in revolutions (error <= 1.3e-7). Their magnitude falls short
// of 1 by ~3e-8 on average, a bias that does not average out over the dense
// bins (it scales every template); *corr receives delta = (1 - c^2 - s^2) / 2,
// formed in FP64 (products of two floats are exact there; in FP32 the sum
// near 1 cannot resolve it), and the caller scales by 1 + delta.
__device__ inline cf cexpi_f32(double phase, float* corr)
{
    double rev = phase * 0.15915494309189535;  // 1 / (2 pi)
    rev -= rint(rev);
    const float r = (float)rev;
    const float c = __builtin_amdgcn_cosf(r), s = __builtin_amdgcn_sinf(r);
    const double cd_ = c, sd_ = s;
    *corr = (float)(-0.5 * fma(cd_, cd_, fma(sd_, sd_, -1.0)));
    return cf{c, s};
}

// add_mode of the FP32-bulk form: t(f) and the phase evaluated in FP64 (the
// same expressions as the FP64 form, so the window decision is identical),
// the other seven cubics and the mode sum in FP32.
__device__ inline void add_mode(cf h[3], const SegRecF& q, double x, double ts, double te, const double2*)
{
    const double x2 = x * x, x3 = x * x2;
    const double tf = q.tf[0] + q.tf[1] * x + q.tf[2] * x2 + q.tf[3] * x3;
    if ((tf >= ts) && ((tf <= te) || (te <= 0.0))) {
        float corr;
        const cf E = cexpi_f32(q.ph[0] + q.ph[1] * x + q.ph[2] * x2 + q.ph[3] * x3, &corr);
        const float xf = (float)x, xf2 = xf * xf, xf3 = xf2 * xf;
        auto ev = [&](int k) { return q.q[k][0] + q.q[k][1] * xf + q.q[k][2] * xf2 + q.q[k][3] * xf3; };
        const float A0 = ev(0);
        const float A = fmaf(A0, corr, A0);  // |e^{i phase}| = 1 restored
        const float ar = A * E.re, ai = A * E.im;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float tr = ev(1 + 2 * c), ti = ev(2 + 2 * c);
            // conj(T A e^{i phase})
            h[c].re += tr * ar - ti * ai;
            h[c].im -= tr * ai + ti * ar;
        }
    }
}

// mode_per_lane of the FP32-bulk form: the whole 192-B record in one round of loads
__device__ inline void mode_per_lane(cf h[3], const SegRecF* __restrict__ rm, int M, int L, double f, int i,
                                     double ts, double te, const double2*)
{
    SegRecF rec = rm[(long)i * M];
    // keeps the compiler from sinking these loads below the bounds check
#pragma unroll
    for (int k = 0; k < 4; k++) asm volatile("" ::"v"(rec.ph[k]), "v"(rec.tf[k]));
#pragma unroll
    for (int k = 0; k < 7; k++) asm volatile("" ::"v"(rec.q[k][0]), "v"(rec.q[k][1]), "v"(rec.q[k][2]), "v"(rec.q[k][3]));
    if (!(rec.x_lo <= f && f < rec.x_hi)) {
        double2 bnd = make_double2(rec.x_lo, rec.x_hi);
        while (f < bnd.x && i > 0) bnd = *reinterpret_cast<const double2*>(rm + (long)--i * M);
        while (f >= bnd.y && i < L - 2) bnd = *reinterpret_cast<const double2*>(rm + (long)++i * M);
        if (!(bnd.x <= f && f < bnd.y)) return;
        rec = rm[(long)i * M];
    }
    add_mode(h, rec, f - rec.x_lo, ts, te, nullptr);
}

constexpr int RB = 2;  // iterations of 64 dense bins per block of a wavefront

// ---------------------------------------------------------------- adjoint helpers

// Adjoint record of one (knot, mode): for each of the nine quantities, the
// adjoints of its segment's (y, c1, c2, c3); after k_dirg_spline_adj, slot 0
// of each quantity holds the adjoint of the quantity's value at the knot.
constexpr int NA = NQ * 4;  // 36 doubles

// quantity of adjoint j (0..7): A, phase, then the six transfer-function parts
// (t_f carries none: the observation window is held fixed)
__device__ inline int adj_q(int j) { return j < 2 ? j : j + 1; }

// d logL with respect to the 32 segment coefficients (A, phase, T1..T3 re/im;
// y, c1, c2, c3 each) of one mode at one dense bin, given
// Lam_c = 4 w_c (d~_c - h~_c) (d logL = Re sum_c conj(Lam_c) dh_c with
// h_c = sum_m conj(T_c A e^{i phase}); Likelihood.cu:357-375 differentiated).
template <class Q>
__device__ inline bool mode_adjoint(double cb[32], const Q& q, double x, double ts, double te, const cd Lam[3],
                                    const double2* tab)
{
    const double x2 = x * x, x3 = x * x2;
    auto ev = [&](int k) {
        const Cubic c = cubic_of(q, k);
        return c.y + c.c1 * x + c.c2 * x2 + c.c3 * x3;
    };
    const double tf = ev(2);
    if (!((tf >= ts) && ((tf <= te) || (te <= 0.0)))) return false;
    const double A = ev(0);
    const cd E = cexpi_tab(ev(1), tab);
    const cd AE = A * E;
    double a[8];
    cd P = mk(0.0, 0.0);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const cd T = mk(ev(3 + 2 * c), ev(4 + 2 * c));
        P += Lam[c] * T;
        const cd LAE = Lam[c] * AE;
        a[2 + 2 * c] = LAE.re;
        a[3 + 2 * c] = -LAE.im;
    }
    const cd PE = P * E;
    a[0] = PE.re;
    a[1] = -A * PE.im;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        cb[4 * j] = a[j];
        cb[4 * j + 1] = a[j] * x;
        cb[4 * j + 2] = a[j] * x2;
        cb[4 * j + 3] = a[j] * x3;
    }
    return true;
}

// One step of reduce32: lanes whose bit BIT is set keep the upper H values,
// the others the lower H, and each adds what its partner (lane ^ BIT) sends.
template <int H, int BIT>
__device__ inline void halve(double v[32], int lane)
{
    const bool up = (lane & BIT) != 0;
#pragma unroll
    for (int j = 0; j < H; j++) {
        const double send = up ? v[j] : v[j + H];
        const double keep = up ? v[j + H] : v[j];
        v[j] = keep + __shfl_xor(send, BIT, 64);
    }
}

// Sum of v[0..32) over the 64 lanes of a wavefront by recursive halving (each
// step exchanges half of the remaining values with the partner lane): lane l
// returns the total of value l >> 1 (31 exchanges instead of 32 x 6).
__device__ inline double reduce32(double v[32], int lane)
{
    halve<16, 32>(v, lane);
    halve<8, 16>(v, lane);
    halve<4, 8>(v, lane);
    halve<2, 4>(v, lane);
    halve<1, 2>(v, lane);
    return v[0] + __shfl_xor(v[0], 1, 64);
}

// adjoint of one lane's bin for one mode, segment searched by the lane (as
// mode_per_lane), added to the segment's adjoint record
__device__ inline void mode_adjoint_per_lane(double* __restrict__ Ym, const SegRec* __restrict__ rm, int M, int L,
                                             double f, int i, double ts, double te, const cd Lam[3],
                                             const double2* tab)
{
    double2 bnd = *reinterpret_cast<const double2*>(rm + (long)i * M);
    while (f < bnd.x && i > 0) bnd = *reinterpret_cast<const double2*>(rm + (long)--i * M);
    while (f >= bnd.y && i < L - 2) bnd = *reinterpret_cast<const double2*>(rm + (long)++i * M);
    if (!(bnd.x <= f && f < bnd.y)) return;
    double cb[32];
    if (!mode_adjoint(cb, *(rm + (long)i * M), f - bnd.x, ts, te, Lam, tab)) return;
    double* y = Ym + (long)i * M * NA;
#pragma unroll
    for (int j = 0; j < 32; j++) unsafeAtomicAdd(y + adj_q(j >> 2) * 4 + (j & 3), cb[j]);
}

// Persistent dense-band kernel: partial[it] = (Re, Im of sum conj(d~) h~, sum |h~|^2, 0)
// over the dense bins of work item it, where h~ = w h (Likelihood.cu:357-375).
//
// Each wavefront takes a contiguous share of the work item's bins, in blocks of
// RB x 64 consecutive bins (lane l: bins kb + 64 r + l), and sums the modes of a
// block mode by mode. The segment of bin f in a mode's grid is
// searchsorted(x, f, 'right') - 1 (waveformbuild.py:139-156). Dense bins are
// much finer than the sparse grid over most of a band, so the lanes of an
// iteration usually share one or two segments: a wavefront-uniform walker then
// holds the current segment, reads its record once with scalar loads (bounds and
// the nine cubics) and evaluates every lane whose f lies in [x_lo, x_hi), moving
// to the next segment until all lanes are placed. The comparisons are the exact
// FP64 ones against the stored knots, so the segment of every bin is the one
// searchsorted gives. Where an iteration spans more than three segments (low
// frequencies, short segments) each lane searches its own (mode_per_lane).
//
// ADJ: also the adjoint pass of the direct log-likelihood's gradient: after the
// mode sum of a block, Lam = 4 w (d~ - h~) per bin and channel, then the modes
// again, each bin's coefficient adjoints (mode_adjoint) summed over the lanes
// that share a segment (reduce32) and added to the segment's adjoint record.
// FP64 forms: at least 4 wavefronts per SIMD (at most 128 VGPRs), measured
// faster than the 3 that the unconstrained register allocation gives, despite a
// few spilled registers in the rarely taken per-lane path; FP32-bulk form: at
// least 6 (at most 80 VGPRs), measured 6 % faster than its unconstrained 5
template <bool ADJ, class Rec>
__global__ __launch_bounds__(256) __attribute__((
    amdgpu_waves_per_eu(std::is_same<Rec, SegRecF>::value ? 6 : 4, 8))) void k_dir_like_t(double* __restrict__ partial, double* __restrict__ Yadj,
                                                    const int* __restrict__ band,
                                                  const int* __restrict__ total, const int* __restrict__ perm,
                                                  int* __restrict__ next, const Rec* __restrict__ R,
                                                  const double* __restrict__ grid_lo,
                                                  const double* __restrict__ grid_istep,
                                                  const long* __restrict__ off, const int* __restrict__ len,
                                                  const double* __restrict__ t_start,
                                                  const double* __restrict__ t_end, const double* __restrict__ freqs,
                                                  const cd* __restrict__ dch, const double* __restrict__ wts, int N,
                                                  int B, int M, int chunk)
{
    static_assert(!ADJ || std::is_same<Rec, SegRec>::value, "the adjoint pass reads FP64 records");
    using HT = typename std::conditional<std::is_same<Rec, SegRecF>::value, cf, cd>::type;
    __shared__ double red[3][256 / 64];
    __shared__ int s_next;
    __shared__ double2 s_tab[129];
    const int nitems = total[0];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
    if constexpr (!std::is_same<Rec, SegRecF>::value) fill_etab(s_tab, tid, blockDim.x);  // before the first barrier
    for (;;) {
        // the next item in the chunk-major order (dynamic: the items running at
        // any time stay a window of consecutive entries, i.e. the same chunks)
        if (tid == 0) s_next = atomicAdd(next, 1);
        __syncthreads();
        const int pi = s_next;
        __syncthreads();
        if (pi >= nitems) break;
        const int it = perm[pi];
        int b;
        const int g = item_chunk(band, B, it, chunk, b);
        const int st = band[3 * b];
        const int k0 = max(st, g * chunk);
        const int k1 = min(band[3 * b + 1], (g + 1) * chunk);
        const long base = off[b];
        const int L = len[b];
        const double ts = t_start[b], te = t_end[b];
        // this wavefront's share: a multiple of 64 bins
        const int per = (((k1 - k0 + nw - 1) / nw) + 63) & ~63;
        const int kw0 = k0 + wv * per, kw1 = min(k1, kw0 + per);
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int kb = kw0; kb < kw1; kb += RB * 64) {
            double f[RB], lf[RB];
            HT h[RB][3];
#pragma unroll
            for (int r = 0; r < RB; r++) {
                const int k = kb + 64 * r + lane;
                f[r] = k < kw1 ? freqs[k] : 0.0;
                lf[r] = (double)__log2f((float)f[r]);
#pragma unroll
                for (int ch = 0; ch < 3; ch++) zero(h[r][ch]);
            }
            for (int m = 0; m < M; m++) {
                const double gis = grid_istep[b * M + m];
                if (!(gis > 0.0)) continue;  // degenerate grid: no segment holds any f
                const double glo = grid_lo[b * M + m];
                const Rec* rm = R + base * M + m;
                int cur = -1;  // the walker's segment (wavefront-uniform); -1: not placed
#pragma unroll
                for (int r = 0; r < RB; r++) {
                    bool pending = kb + 64 * r + lane < kw1;
                    const unsigned long long pm = __ballot(pending);
                    if (pm == 0) break;  // later iterations are past the share too
                    const int est = (int)fmin(fmax((lf[r] - glo) * gis, 0.0), (double)(L - 2));
                    const int e_first = __builtin_amdgcn_readlane(est, __builtin_ctzll(pm));
                    const int e_last = __builtin_amdgcn_readlane(est, 63 - __builtin_clzll(pm));
                    if (e_last - e_first > 2) {
                        if (pending) mode_per_lane(h[r], rm, M, L, f[r], est, ts, te, s_tab);
                        cur = -1;
                        continue;
                    }
                    if (cur < 0) cur = e_first;
                    // the walker moves down only while a pending lane lies below the current
                    // segment, up only past lanes it has placed: in a sorted grid every step
                    // makes progress; the bound only guards against an unsorted one
                    for (int guard = 2 * L + 8; guard > 0; guard--) {
                        if (__ballot(pending) == 0) break;
                        const Rec* rc = rm + (long)cur * M;
                        const double xlo = rc->x_lo, xhi = rc->x_hi;
                        const bool below = pending && f[r] < xlo;
                        if (__ballot(below) != 0) {
                            if (cur == 0) pending = pending && !below;  // left of the grid
                            else cur--;
                            continue;
                        }
                        const bool in = pending && f[r] < xhi;
                        if (in) add_mode(h[r], *rc, f[r] - xlo, ts, te, s_tab);
                        pending = pending && !in;
                        if (__ballot(pending) != 0) {
                            if (cur == L - 2) pending = false;  // right of the grid
                            else cur++;
                        }
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < RB; r++) {
                const int k = kb + 64 * r + lane;
                if (k < kw1) {
#pragma unroll
                    for (int ch = 0; ch < 3; ch++) {
                        const double w = wts[(long)ch * N + k];
                        const cd ht = to_cd(h[r][ch]) * w;
                        const cd d = dch[(long)ch * N + k];
                        const cd dh = conj(d) * ht;
                        a0 += dh.re;
                        a1 += dh.im;
                        a2 += ht.re * ht.re + ht.im * ht.im;
                        if constexpr (ADJ) h[r][ch] = (4.0 * w) * (d - ht);  // Lam
                    }
                }
            }
            if constexpr (ADJ) {
                for (int m = 0; m < M; m++) {
                    const double gis = grid_istep[b * M + m];
                    if (!(gis > 0.0)) continue;
                    const double glo = grid_lo[b * M + m];
                    const SegRec* rm = R + base * M + m;
                    double* Ym = Yadj + (base * M + m) * (long)NA;
                    int cur = -1;
#pragma unroll
                    for (int r = 0; r < RB; r++) {
                        bool pending = kb + 64 * r + lane < kw1;
                        const unsigned long long pm = __ballot(pending);
                        if (pm == 0) break;
                        const int est = (int)fmin(fmax((lf[r] - glo) * gis, 0.0), (double)(L - 2));
                        const int e_first = __builtin_amdgcn_readlane(est, __builtin_ctzll(pm));
                        const int e_last = __builtin_amdgcn_readlane(est, 63 - __builtin_clzll(pm));
                        if (e_last - e_first > 2) {
                            if (pending) mode_adjoint_per_lane(Ym, rm, M, L, f[r], est, ts, te, h[r], s_tab);
                            cur = -1;
                            continue;
                        }
                        if (cur < 0) cur = e_first;
                        for (int guard = 2 * L + 8; guard > 0; guard--) {
                            if (__ballot(pending) == 0) break;
                            const SegRec* rc = rm + (long)cur * M;
                            const double xlo = rc->x_lo, xhi = rc->x_hi;
                            const bool below = pending && f[r] < xlo;
                            if (__ballot(below) != 0) {
                                if (cur == 0) pending = pending && !below;
                                else cur--;
                                continue;
                            }
                            const bool in = pending && f[r] < xhi;
                            double cb[32];
#pragma unroll
                            for (int j = 0; j < 32; j++) cb[j] = 0.0;
                            bool used = false;
                            if (in) used = mode_adjoint(cb, *rc, f[r] - xlo, ts, te, h[r], s_tab);
                            if (__ballot(used) != 0) {
                                const double tot = reduce32(cb, lane);
                                if ((lane & 1) == 0) {
                                    const int j = lane >> 1;
                                    unsafeAtomicAdd(Ym + (long)cur * M * NA + adj_q(j >> 2) * 4 + (j & 3), tot);
                                }
                            }
                            pending = pending && !in;
                            if (__ballot(pending) != 0) {
                                if (cur == L - 2) pending = false;
                                else cur++;
                            }
                        }
                    }
                }
            }
        }
        // block reduction in a fixed order
        a0 = wsum(a0);
        a1 = wsum(a1);
        a2 = wsum(a2);
        if (lane == 0) {
            red[0][wv] = a0;
            red[1][wv] = a1;
            red[2][wv] = a2;
        }
        __syncthreads();
        if (tid == 0) {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
            for (int w = 0; w < nw; w++) {
                s0 += red[0][w];
                s1 += red[1][w];
                s2 += red[2][w];
            }
            partial[4 * (long)it] = s0;
            partial[4 * (long)it + 1] = s1;
            partial[4 * (long)it + 2] = s2;
            partial[4 * (long)it + 3] = 0.0;
        }
        __syncthreads();
    }
}

// One wavefront per binary: lane l adds the binary's items l, l + 64, ... (in
// item order), then the wavefront adds the 64 lane sums in a fixed tree; the
// result does not depend on the schedule of the dense-band kernel.
__global__ __launch_bounds__(64) void k_dir_reduce(double* __restrict__ d_h, double* __restrict__ h_h,
                                                   const double* __restrict__ partial, const int* __restrict__ band,
                                                   const int* __restrict__ total, int B)
{
    const int b = blockIdx.x;
    const int c0 = band[3 * b + 2];
    const int c1 = b + 1 < B ? band[3 * (b + 1) + 2] : total[0];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int c = c0 + (int)threadIdx.x; c < c1; c += 64) {
        s0 += partial[4 * (long)c];
        s1 += partial[4 * (long)c + 1];
        s2 += partial[4 * (long)c + 2];
    }
    s0 = wsum(s0);
    s1 = wsum(s1);
    s2 = wsum(s2);
    if (threadIdx.x == 0) {
        d_h[2 * b] = 4.0 * s0;
        d_h[2 * b + 1] = 4.0 * s1;
        h_h[2 * b] = 4.0 * s2;
        h_h[2 * b + 1] = 0.0;
    }
}

// ---------------------------------------------------------------- FP32-bulk routing
//
// The FP32-bulk form's error, measured template by template against the FP64
// form on the uneven (8192) and the large (512) batches with every template in
// FP32, is at most 1.4e-8 h_h wherever h_h > 3000 (and below 3.3e-5 in logL
// where h_h <= 3000), so templates with h_h above FP32_FALLBACK_HH (error up to
// ~3e-4 in logL there, 3x inside the 1e-3 target) are evaluated with the FP64
// records and kernel. Which ones is first
// decided from an estimate of h_h made before the dense pass (k_dir_route);
// after it, any binary the estimate sent to FP32 whose h_h exceeds the bound
// is evaluated again in FP64 (its work items only), so the estimate decides
// the cost and never the result.

constexpr double FP32_FALLBACK_HH = 2.0e4;

__global__ void k_dir_loud_zero(int* __restrict__ loud_cnt)
{
    if (threadIdx.x < 4) loud_cnt[threadIdx.x] = 0;
}

// h_h of each binary estimated from its sparse knots (one wavefront per
// binary), each mode on its own grid and the modes added without their
// interference:
//   4 sum_m sum_c sum_i A_mi^2 |T_mci|^2 w_c(x_i)^2 (x_{i+1} - x_i) / df
// over knots inside the data band and the observation window, w taken at the
// nearest bin of a uniform dense grid (for another grid the estimate, and so
// only the split between the passes, is coarser).
__global__ __launch_bounds__(64) void k_dir_route(int* __restrict__ route, const double* __restrict__ X,
                                                  const double* __restrict__ Yc, const long* __restrict__ off,
                                                  const int* __restrict__ len, const double* __restrict__ t_start,
                                                  const double* __restrict__ t_end, const double* __restrict__ freqs,
                                                  const double* __restrict__ wts, int N, int M)
{
    const int b = blockIdx.x, lane = threadIdx.x;
    const int L = len[b];
    const long base = off[b];
    const double ts = t_start[b], te = t_end[b];
    const double f0 = freqs[0], f1 = freqs[N - 1];
    const double idf = f1 > f0 ? (N - 1) / (f1 - f0) : 0.0;
    double s = 0.0;
    // entry j = i M + m (knot i of mode m): consecutive lanes read consecutive
    // (knot, mode) entries of X and Yc
    for (int j = lane; j < (L - 1) * M; j += 64) {
        const long p = base * M + j;
        const double x = X[p], xn = X[p + M];
        if (!(x >= f0 && x <= f1 && xn > x)) continue;
        const double* y = Yc + p * NQ;
        const double tf = y[2];
        if (!((tf >= ts) && ((tf <= te) || (te <= 0.0)))) continue;
        const int k = min(N - 1, max(0, (int)((x - f0) * idf + 0.5)));
        double t = 0.0;
        for (int ch = 0; ch < 3; ch++) {
            const double w = wts[(long)ch * N + k];
            t += (y[3 + 2 * ch] * y[3 + 2 * ch] + y[4 + 2 * ch] * y[4 + 2 * ch]) * (w * w);
        }
        s += (y[0] * y[0]) * t * ((fmin(xn, f1) - x) * idf);
    }
    s = wsum(s);
    if (lane == 0) route[b] = 4.0 * s > FP32_FALLBACK_HH ? 1 : 0;
}

// The chunk-major order split into the items of binaries routed to FP64
// (`loud`) and the others (`quiet`), each list keeping the order (the items in
// flight then share their dense chunks, whose data come from L2): a count of
// loud entries per block of SPLIT_BLOCK entries, its exclusive scan
// (k_dir_cnt_scan), and the placement of every entry.
constexpr int SPLIT_BLOCK = 256;

__device__ inline bool routed_loud(int pi, int n, const int* __restrict__ perm, const int* __restrict__ band, int B,
                                   const int* __restrict__ route, int& it)
{
    it = pi < n ? perm[pi] : 0;
    return pi < n && route[owner(band, B, it)] != 0;
}

__global__ __launch_bounds__(SPLIT_BLOCK) void k_dir_split_count(int* __restrict__ split_cnt,
                                                                 const int* __restrict__ perm,
                                                                 const int* __restrict__ total,
                                                                 const int* __restrict__ band, int B,
                                                                 const int* __restrict__ route)
{
    __shared__ int wc[SPLIT_BLOCK / 64];
    int it;
    const bool l = routed_loud(blockIdx.x * SPLIT_BLOCK + threadIdx.x, total[0], perm, band, B, route, it);
    const unsigned long long m = __ballot(l);
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int c = 0;
        for (int w = 0; w < SPLIT_BLOCK / 64; w++) c += wc[w];
        split_cnt[blockIdx.x] = c;
    }
}

// after the scan: split_cnt[blk] = loud entries before block blk. Entry pi goes
// to loud[#loud before pi] or quiet[pi - #loud before pi]; the last entry
// writes both counts (cnt[0] loud, cnt[2] quiet; zeroed before, for an empty order)
__global__ __launch_bounds__(SPLIT_BLOCK) void k_dir_split_write(int* __restrict__ loud, int* __restrict__ quiet,
                                                                 int* __restrict__ cnt,
                                                                 const int* __restrict__ split_cnt,
                                                                 const int* __restrict__ perm,
                                                                 const int* __restrict__ total,
                                                                 const int* __restrict__ band, int B,
                                                                 const int* __restrict__ route)
{
    __shared__ int wc[SPLIT_BLOCK / 64];
    const int n = total[0];
    const int pi = blockIdx.x * SPLIT_BLOCK + threadIdx.x;
    int it;
    const bool l = routed_loud(pi, n, perm, band, B, route, it);
    const unsigned long long m = __ballot(l);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wc[w] = __popcll(m);
    __syncthreads();
    int before = split_cnt[blockIdx.x];
    for (int k = 0; k < w; k++) before += wc[k];
    const int lpos = before + __popcll(m & ((1ull << lane) - 1ull));
    if (pi < n) {
        if (l) loud[lpos] = it;
        else quiet[pi - lpos] = it;
        if (pi == n - 1) {
            cnt[0] = lpos + (l ? 1 : 0);
            cnt[2] = n - cnt[0];
        }
    }
}

// after the passes: the items of binaries evaluated in FP32 whose h_h exceeds
// the bound
__global__ void k_dir_loud_list(int* __restrict__ loud, int* __restrict__ loud_cnt, const int* __restrict__ perm,
                                const int* __restrict__ total, const int* __restrict__ band, int B,
                                const double* __restrict__ h_h, const int* __restrict__ route)
{
    const int pi = blockIdx.x * blockDim.x + threadIdx.x;
    if (pi >= total[0]) return;
    const int it = perm[pi];
    const int b = owner(band, B, it);
    if (!route[b] && h_h[2 * b] > FP32_FALLBACK_HH) loud[atomicAdd(loud_cnt, 1)] = it;
}

// ---------------------------------------------------------------- gradient kernels

__global__ void k_dirg_zero(double* __restrict__ y, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = 0.0;
}

// records with tangents, one lane per (binary, mode)
__global__ __launch_bounds__(64) void k_dirg_setup(gradk::GradRec* __restrict__ grecs,
                                                   const double* __restrict__ params, int B, ModeList modes, int M,
                                                   const double* __restrict__ qnm, int table_len, double dspin)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * M) return;
    const int b = idx / M, mi = idx % M;
    gradk::grad_setup_one(grecs[idx], params, B, b, modes.ell[mi], modes.mm[mi], qnm, table_len, dspin);
}

// Adjoint of the not-a-knot spline construction (the system of k_dir_spline):
// segment-coefficient adjoints -> adjoints of the knot values. One wavefront
// per binary, one lane per (mode, quantity) system (t_f excluded). In the
// record of knot i, quantity q: slots 0..3 hold the coefficient adjoints of
// segment i on entry; the passes reuse them (0: value adjoint, 1: slope
// adjoint then z, 2, 3: the transposed solve's c', d'); slot 0 holds the
// knot-value adjoint on exit.
__global__ __launch_bounds__(64) void k_dirg_spline_adj(double* __restrict__ Yadj, const double* __restrict__ X,
                                                        const long* __restrict__ off, const int* __restrict__ len,
                                                        int M)
{
    const int b = blockIdx.x;
    const int s = threadIdx.x;
    if (s >= M * 8) return;
    const int m = s / 8, q = adj_q(s % 8);
    const int n = len[b];
    const long base = off[b];
    auto xi = [&](int i) { return X[(base + i) * M + m]; };
    auto Y = [&](int i) { return Yadj + ((base + i) * M + m) * (long)NA + q * 4; };
    // pass 1: segment coefficients -> value and slope adjoints (carry into knot i + 1)
    double cy = 0.0, cs = 0.0;
    for (int i = 0; i < n - 1; i++) {
        double* y = Y(i);
        const double c0 = y[0], c1 = y[1], c2 = y[2], c3 = y[3];
        const double dx = xi(i + 1) - xi(i);
        const double tb = -c2 + c3 / dx;
        const double slb = c2 / dx - 2.0 * tb / dx;
        y[0] = cy + c0 - slb / dx;
        y[1] = cs + c1 - c2 / dx + tb / dx;
        cy = slb / dx;
        cs = tb / dx;
    }
    {
        double* y = Y(n - 1);
        y[0] = cy;
        y[1] = cs;
    }
    // pass 2: A^T z = slope adjoints (A: the tridiagonal matrix of the slopes)
    auto Aij = [&](int i, int j) -> double {
        if (i == 0) return j == 0 ? xi(2) - xi(1) : (j == 1 ? xi(2) - xi(0) : 0.0);
        if (i == n - 1) return j == n - 2 ? xi(n - 1) - xi(n - 3) : (j == n - 1 ? xi(n - 2) - xi(n - 3) : 0.0);
        if (j == i - 1) return xi(i + 1) - xi(i);
        if (j == i) return 2.0 * (xi(i + 1) - xi(i - 1));
        if (j == i + 1) return xi(i) - xi(i - 1);
        return 0.0;
    };
    double cprev = 0.0, dprev = 0.0;
    for (int i = 0; i < n; i++) {
        double* y = Y(i);
        const double bdiag = Aij(i, i);
        double den, d;
        if (i == 0) {
            den = bdiag;
            d = y[1] / den;
        } else {
            const double a = Aij(i - 1, i);
            den = bdiag - a * cprev;
            d = (y[1] - a * dprev) / den;
        }
        const double c = i < n - 1 ? Aij(i + 1, i) / den : 0.0;
        y[2] = c;
        y[3] = d;
        cprev = c;
        dprev = d;
    }
    double znext = Y(n - 1)[3];
    Y(n - 1)[1] = znext;
    for (int i = n - 2; i >= 0; i--) {
        double* y = Y(i);
        znext = y[3] - y[2] * znext;
        y[1] = znext;
    }
    // pass 3: value adjoints += R^T z (R: the right-hand side's dependence on the values)
    {
        const double dx1 = xi(1) - xi(0), dx2 = xi(2) - xi(1), dd = xi(2) - xi(0);
        const double a1 = (dx1 + 2.0 * dd) * dx2 / (dd * dx1), a2 = dx1 * dx1 / (dd * dx2);
        const double z0 = Y(0)[1];
        Y(0)[0] += -a1 * z0;
        Y(1)[0] += (a1 - a2) * z0;
        Y(2)[0] += a2 * z0;
    }
    for (int i = 1; i < n - 1; i++) {
        const double e1 = xi(i) - xi(i - 1), e2 = xi(i + 1) - xi(i);
        const double b1 = 3.0 * e2 / e1, b2 = 3.0 * e1 / e2;
        const double zi = Y(i)[1];
        Y(i - 1)[0] += -b1 * zi;
        Y(i)[0] += (b1 - b2) * zi;
        Y(i + 1)[0] += b2 * zi;
    }
    {
        const double e1 = xi(n - 2) - xi(n - 3), e2 = xi(n - 1) - xi(n - 2), ed = xi(n - 1) - xi(n - 3);
        const double c1 = e2 * e2 / (ed * e1), c2 = (2.0 * ed + e2) * e1 / (ed * e2);
        const double zn = Y(n - 1)[1];
        Y(n - 3)[0] += -c1 * zn;
        Y(n - 2)[0] += (c1 - c2) * zn;
        Y(n - 1)[0] += c2 * zn;
    }
}

// Knot-level contraction: one wavefront per (binary, mode), lane per knot; the
// knot-value adjoints of A, phase and T weight the waveform's and response's
// tangents at the knot's frequency (grad_point_core, no window rule: the
// window acts on the dense bins). partial[(b M + m) 11 + p].
__global__ __launch_bounds__(64) void k_dirg_knots(double* __restrict__ gpart, const gradk::GradRec* __restrict__ grecs,
                                                   const double* __restrict__ Yadj, const double* __restrict__ X,
                                                   const long* __restrict__ off, const int* __restrict__ len,
                                                   const double* __restrict__ params, int B, int M,
                                                   resp::OrbitTab orb, int TDItag, int rescaled, int tdi2,
                                                   double initial_t_val)
{
    const int bm = blockIdx.x;
    const int b = bm / M, m = bm % M;
    Window win;
    win.t_ref = params[P_TREF * B + b];
    win.ts = 0.0;
    win.te = 0.0;
    win.dt_shift = win.t_ref - initial_t_val;
    const double dist = params[P_DIST * B + b];
    const gradk::GradRec& G = grecs[bm];
    const int n = len[b];
    const long base = off[b];
    __shared__ double2 tab[129];
    etab_to_lds(tab);
    __syncthreads();
    double acc[gradk::NGRAD];
    for (int p = 0; p < gradk::NGRAD; p++) acc[p] = 0.0;
    for (int i = threadIdx.x; i < n; i += 64) {
        const long rec = (base + i) * M + m;
        const double* y = Yadj + rec * NA;
        const double aA = y[0 * 4], aPh = y[1 * 4];
        cd Wt[3];
        for (int c = 0; c < 3; c++) Wt[c] = mk(y[(3 + 2 * c) * 4], -y[(4 + 2 * c) * 4]);
        auto weights = [&](const cd*, cd, double, cd W[3], double& a_A, double& a_Ph) {
            for (int c = 0; c < 3; c++) W[c] = Wt[c];
            a_A = aA;
            a_Ph = aPh;
        };
        gradk::grad_point_core(G, orb, X[rec], win, false, weights, TDItag, rescaled, tdi2, dist, acc, tab);
    }
    for (int p = 0; p < gradk::NGRAD; p++) acc[p] = wsum(acc[p]);
    if (threadIdx.x == 0)
        for (int p = 0; p < gradk::NGRAD; p++) gpart[(long)bm * gradk::NGRAD + p] = acc[p];
}

// grad[p B + b] = sum over modes, in mode order
__global__ void k_dirg_reduce(double* __restrict__ grad, const double* __restrict__ gpart, int B, int M)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * gradk::NGRAD) return;
    const int p = t / B, b = t % B;
    double s = 0.0;
    for (int m = 0; m < M; m++) s += gpart[((long)b * M + m) * gradk::NGRAD + p];
    grad[(long)p * B + b] = s;
}

template <class T>
T* dalloc(size_t n, bool& ok)
{
    void* p = nullptr;
    if (!ok || hipMalloc(&p, n * sizeof(T) + 256) != hipSuccess) {
        ok = false;
        return nullptr;
    }
    return static_cast<T*>(p);
}

void plan_free(bbhx_direct_plan* p)
{
    void* bufs[] = {p->off, p->len, p->recs, p->X, p->Cf, p->Yc, p->Fs, p->grid_lo, p->grid_istep, p->band,
                    p->total, p->partial, p->perm, p->cnt, p->next, p->Yadj, p->grecs, p->gpart, p->CfF,
                    p->loud, p->quiet, p->loud_cnt, p->route, p->split_cnt};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
}

}  // namespace

extern "C" int bbhx_hip_direct_plan_create(bbhx_direct_plan** out, int numBinAll, const int* lengths_host,
                                           int numModes, int data_length, int min_chunk)
{
    if (!out || numBinAll < 1 || !lengths_host || numModes < 1 || numModes > MAX_MODES || data_length < 1 ||
        min_chunk < 64)
        return ERR_ARGS;
    std::vector<long> off(numBinAll);
    long S = 0;
    int Lmax = 0;
    for (int b = 0; b < numBinAll; b++) {
        if (lengths_host[b] < 4) return ERR_ARGS;
        off[b] = S;
        S += lengths_host[b];
        Lmax = lengths_host[b] > Lmax ? lengths_host[b] : Lmax;
    }
    // the default-grid t(f) of one (binary, mode) lives in LDS
    if ((size_t)Lmax * sizeof(double) > 48 * 1024) return ERR_ARGS;
    auto* p = new bbhx_direct_plan;
    (void)hipGetDevice(&p->device);
    p->B = numBinAll;
    p->M = numModes;
    p->N = data_length;
    p->Lmax = Lmax;
    p->min_chunk = min_chunk;
    p->S = S;
    p->max_items = (long)numBinAll * ((data_length + min_chunk - 1) / min_chunk) + numBinAll;
    if (p->max_items > 2147483647L) {
        delete p;
        return ERR_ARGS;
    }
    bool ok = true;
    const size_t SM = (size_t)S * numModes;
    p->off = dalloc<long>(numBinAll, ok);
    p->len = dalloc<int>(numBinAll, ok);
    p->recs = dalloc<FusedRec>((size_t)numBinAll * numModes, ok);
    p->X = dalloc<double>(SM, ok);
    p->Cf = dalloc<double>(SM * RS, ok);
    p->Yc = dalloc<double>(SM * NQ, ok);
    p->Fs = dalloc<double>(SM * (NQ + 1), ok);
    p->grid_lo = dalloc<double>((size_t)numBinAll * numModes, ok);
    p->grid_istep = dalloc<double>((size_t)numBinAll * numModes, ok);
    p->band = dalloc<int>((size_t)numBinAll * 3, ok);
    p->total = dalloc<int>(1, ok);
    p->partial = dalloc<double>((size_t)p->max_items * 4, ok);
    p->max_chunks = (data_length + min_chunk - 1) / min_chunk;
    p->perm = dalloc<int>((size_t)p->max_items, ok);
    p->cnt = dalloc<int>((size_t)p->max_chunks + 1, ok);
    p->next = dalloc<int>(1, ok);
    if (!ok) {
        plan_free(p);
        delete p;
        return ERR_ALLOC;
    }
    if (hipMemcpy(p->off, off.data(), sizeof(long) * numBinAll, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->len, lengths_host, sizeof(int) * numBinAll, hipMemcpyHostToDevice) != hipSuccess) {
        plan_free(p);
        delete p;
        return ERR_LAUNCH;
    }
    *out = p;
    return OK;
}

extern "C" int bbhx_hip_direct_plan_destroy(bbhx_direct_plan* p)
{
    if (!p) return ERR_ARGS;
    (void)hipDeviceSynchronize();
    plan_free(p);
    delete p;
    return OK;
}

extern "C" long bbhx_hip_direct_plan_bytes(const bbhx_direct_plan* p)
{
    if (!p) return -1;
    const long SM = p->S * p->M;
    long bytes = SM * 8 * (1 + RS + NQ + NQ + 1) + (long)p->B * p->M * (sizeof(FusedRec) + 16) + p->max_items * 36 + p->B * 24 +
                 4L * (p->max_chunks + 1);
    if (p->Yadj) bytes += SM * 8 * NA + (long)p->B * p->M * (sizeof(bbhx::gradk::GradRec) + 8 * bbhx::gradk::NGRAD);
    if (p->CfF) bytes += SM * (long)sizeof(SegRecF) + p->max_items * 8 + 16 + p->B * 4L + 4L * (p->max_items / 256 + 1);
    return bytes;
}

extern "C" int bbhx_hip_direct_plan_band(const bbhx_direct_plan* p, int* band_host)
{
    if (!p || !band_host) return ERR_ARGS;
    std::vector<int> b3((size_t)p->B * 3);
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(b3.data(), p->band, sizeof(int) * b3.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return ERR_LAUNCH;
    for (int b = 0; b < p->B; b++) {
        band_host[2 * b] = b3[3 * b];
        band_host[2 * b + 1] = b3[3 * b + 1];
    }
    return OK;
}

namespace {

// The launch sequence of the fused direct likelihood; with grad != nullptr
// also its gradient (the plan's gradient buffers must exist).
int direct_launch(bbhx_direct_plan* p, double* d_h, double* h_h, double* grad, int* flags, const double* params,
                  const double* t_start, const double* t_end, const int* ells_host, const int* mms_host,
                  int numModes, const double* dataFreqs, const double* dataChannels, const double* noiseFactors,
                  const double* qnm, int table_len, double dspin, double initial_t_val, double mf_min, double mf_max,
                  const bbhx_response* orbits, int TDItag, int rescaled, int tdi2, int chunk, int block, int grid,
                  void* stream)
{
    if (!p) return ERR_ARGS;
    if (chunk < p->min_chunk || (block != 64 && block != 128 && block != 256) || grid < 1) return ERR_ARGS;
    int dev = -1;
    (void)hipGetDevice(&dev);
    if (dev != p->device || numModes != p->M) return ERR_ARGS;
    resp::OrbitTab orb;
    if (!response_orbit_tab(orbits, orb)) return ERR_ARGS;
    ModeList ml;
    const int mst = check_modes(ells_host, mms_host, numModes, ml);
    if (mst != OK) return mst;
    if (!d_h || !h_h || !flags || !params || !t_start || !t_end || !dataFreqs || !dataChannels || !noiseFactors ||
        !qnm || table_len < 2 || !(mf_min > 0.0) || !(mf_max > mf_min))
        return ERR_ARGS;
    if (grad && !p->Yadj) return ERR_ARGS;  // bbhx_hip_direct_plan_enable_grad first
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int B = p->B, M = p->M;
    hipEvent_t ev[3];
    const bool timed = timer_take3(ev);
    if (timed) (void)hipEventRecord(ev[0], st);
    hipLaunchKernelGGL(k_dir_setup, dim3(setup_blocks(B, M)), dim3(64), 0, st, p->recs, params, B,
                       ml, M, qnm,
                       table_len, dspin);
    hipLaunchKernelGGL(k_dir_grid, dim3(B * M), dim3(64), sizeof(double) * (size_t)p->Lmax, st, p->X, p->Yc,
                       p->grid_lo, p->grid_istep, flags, p->recs, p->off, p->len, params, t_start, t_end, B, M, ml,
                       orb, TDItag, rescaled, tdi2, initial_t_val, mf_min, mf_max);
    // the FP32-bulk form (not for the gradient, whose adjoint pass reads the FP64 records)
    const bool f32 = !grad && p->precision == 1;
    hipLaunchKernelGGL(k_dir_spline, dim3(B), dim3(64), 0, st, p->X, p->Yc, p->Cf, p->Fs,
                       f32 ? static_cast<SegRecF*>(p->CfF) : nullptr, p->off, p->len, M);
    hipLaunchKernelGGL(k_dir_band, dim3(1), dim3(1024), 0, st, p->band, p->total, p->X, p->off, p->len, dataFreqs,
                       p->N, B, M, chunk);
    const int nchunks = (p->N + chunk - 1) / chunk;  // <= max_chunks since chunk >= min_chunk
    hipLaunchKernelGGL(k_dir_cnt_zero, dim3(ceil_div(nchunks + 1, 256)), dim3(256), 0, st, p->cnt, nchunks + 1,
                       p->next);
    // one lane per possible work item (those past the batch's count exit)
    const int item_blocks = ceil_div(p->max_items, 256);
    hipLaunchKernelGGL(k_dir_cnt, dim3(item_blocks), dim3(256), 0, st, p->cnt, p->band, p->total, B, chunk);
    hipLaunchKernelGGL(k_dir_cnt_scan, dim3(1), dim3(1024), 0, st, p->cnt, nchunks);
    hipLaunchKernelGGL(k_dir_perm, dim3(item_blocks), dim3(256), 0, st, p->perm, p->cnt, p->band, p->total, B, chunk);
    if (grad) {
        hipLaunchKernelGGL(k_dirg_zero, dim3(4096), dim3(256), 0, st, p->Yadj, p->S * (long)M * NA);
        hipLaunchKernelGGL(k_dirg_setup, dim3(ceil_div((long)B * M, 64)), dim3(64), 0, st, p->grecs, params, B, ml, M,
                           qnm, table_len, dspin);
    }
    if (timed) (void)hipEventRecord(ev[1], st);
    if (grad)
        hipLaunchKernelGGL((k_dir_like_t<true, SegRec>), dim3(grid), dim3(block), 0, st, p->partial, p->Yadj, p->band, p->total,
                           p->perm, p->next, reinterpret_cast<const SegRec*>(p->Cf), p->grid_lo, p->grid_istep,
                           p->off, p->len, t_start, t_end, dataFreqs, reinterpret_cast<const cd*>(dataChannels),
                           noiseFactors, p->N, B, M, chunk);
    else if (f32) {
        // binaries whose h_h estimate exceeds the bound go to the FP64 pass, the
        // others to the FP32 pass; the FP32-bulk kernel holds 6 wavefronts per SIMD
        // against the FP64 form's 4, so a grid sized for the FP64 form is scaled by
        // 3/2 to fill the chip as fully (measured at 8192 uneven templates, all in
        // FP32: grid 1024 as given 253 ms, 2048 203 ms)
        hipLaunchKernelGGL(k_dir_route, dim3(B), dim3(64), 0, st, p->route, p->X, p->Yc, p->off, p->len, t_start,
                           t_end, dataFreqs, noiseFactors, p->N, M);
        hipLaunchKernelGGL(k_dir_loud_zero, dim3(1), dim3(64), 0, st, p->loud_cnt);
        const int split_blocks = ceil_div(p->max_items, SPLIT_BLOCK);  // <= max_items / 256 + 1
        hipLaunchKernelGGL(k_dir_split_count, dim3(split_blocks), dim3(SPLIT_BLOCK), 0, st, p->split_cnt, p->perm,
                           p->total, p->band, B, p->route);
        hipLaunchKernelGGL(k_dir_cnt_scan, dim3(1), dim3(1024), 0, st, p->split_cnt, split_blocks);
        hipLaunchKernelGGL(k_dir_split_write, dim3(split_blocks), dim3(SPLIT_BLOCK), 0, st, p->loud, p->quiet,
                           p->loud_cnt, p->split_cnt, p->perm, p->total, p->band, B, p->route);
        hipLaunchKernelGGL((k_dir_like_t<false, SegRecF>), dim3((unsigned)std::min<long>(3L * grid / 2, 2147483647L)),
                           dim3(block), 0, st, p->partial, nullptr,
                           p->band, p->loud_cnt + 2, p->quiet, p->loud_cnt + 3, static_cast<const SegRecF*>(p->CfF),
                           p->grid_lo, p->grid_istep, p->off, p->len, t_start, t_end, dataFreqs,
                           reinterpret_cast<const cd*>(dataChannels), noiseFactors, p->N, B, M, chunk);
        hipLaunchKernelGGL((k_dir_like_t<false, SegRec>), dim3(grid), dim3(block), 0, st, p->partial, nullptr,
                           p->band, p->loud_cnt, p->loud, p->loud_cnt + 1, reinterpret_cast<const SegRec*>(p->Cf),
                           p->grid_lo, p->grid_istep, p->off, p->len, t_start, t_end, dataFreqs,
                           reinterpret_cast<const cd*>(dataChannels), noiseFactors, p->N, B, M, chunk);
    } else
        hipLaunchKernelGGL((k_dir_like_t<false, SegRec>), dim3(grid), dim3(block), 0, st, p->partial, nullptr,
                           p->band, p->total, p->perm, p->next, reinterpret_cast<const SegRec*>(p->Cf), p->grid_lo,
                           p->grid_istep, p->off, p->len, t_start, t_end, dataFreqs,
                           reinterpret_cast<const cd*>(dataChannels), noiseFactors, p->N, B, M, chunk);
    hipLaunchKernelGGL(k_dir_reduce, dim3(B), dim3(64), 0, st, d_h, h_h, p->partial, p->band,
                       p->total, B);
    if (f32) {
        // binaries of the FP32 pass whose h_h exceeds the bound again in FP64 (their
        // items' partials overwritten), then every binary reduced again (the others'
        // partials are unchanged)
        hipLaunchKernelGGL(k_dir_loud_zero, dim3(1), dim3(64), 0, st, p->loud_cnt);
        hipLaunchKernelGGL(k_dir_loud_list, dim3(item_blocks), dim3(256), 0, st, p->loud, p->loud_cnt, p->perm,
                           p->total, p->band, B, h_h, p->route);
        hipLaunchKernelGGL((k_dir_like_t<false, SegRec>), dim3(grid), dim3(block), 0, st, p->partial, nullptr,
                           p->band, p->loud_cnt, p->loud, p->loud_cnt + 1, reinterpret_cast<const SegRec*>(p->Cf),
                           p->grid_lo, p->grid_istep, p->off, p->len, t_start, t_end, dataFreqs,
                           reinterpret_cast<const cd*>(dataChannels), noiseFactors, p->N, B, M, chunk);
        hipLaunchKernelGGL(k_dir_reduce, dim3(B), dim3(64), 0, st, d_h, h_h, p->partial, p->band, p->total, B);
    }
    if (grad) {
        hipLaunchKernelGGL(k_dirg_spline_adj, dim3(B), dim3(64), 0, st, p->Yadj, p->X, p->off, p->len, M);
        hipLaunchKernelGGL(k_dirg_knots, dim3(B * M), dim3(64), 0, st, p->gpart, p->grecs, p->Yadj, p->X, p->off,
                           p->len, params, B, M, orb, TDItag, rescaled, tdi2, initial_t_val);
        hipLaunchKernelGGL(k_dirg_reduce, dim3(ceil_div((long)B * gradk::NGRAD, 256)), dim3(256), 0, st, grad,
                           p->gpart, B, M);
    }
    if (timed) (void)hipEventRecord(ev[2], st);
    return hipGetLastError() == hipSuccess ? OK : ERR_LAUNCH;
}

}  // namespace

extern "C" int bbhx_hip_direct_fused(bbhx_direct_plan* p, double* d_h, double* h_h, int* flags, const double* params,
                                     const double* t_start, const double* t_end, const int* ells_host,
                                     const int* mms_host, int numModes, const double* dataFreqs,
                                     const double* dataChannels, const double* noiseFactors, const double* qnm,
                                     int table_len, double dspin, double initial_t_val, double mf_min, double mf_max,
                                     const bbhx_response* orbits, int TDItag, int rescaled, int tdi2, int chunk,
                                     int block, int grid, void* stream)
{
    return direct_launch(p, d_h, h_h, nullptr, flags, params, t_start, t_end, ells_host, mms_host, numModes,
                         dataFreqs, dataChannels, noiseFactors, qnm, table_len, dspin, initial_t_val, mf_min, mf_max,
                         orbits, TDItag, rescaled, tdi2, chunk, block, grid, stream);
}

extern "C" int bbhx_hip_direct_plan_set_precision(bbhx_direct_plan* p, int precision)
{
    if (!p || (precision != 0 && precision != 1)) return ERR_ARGS;
    if (precision == 1 && !p->CfF) {
        bool ok = true;
        p->CfF = dalloc<SegRecF>((size_t)p->S * p->M, ok);
        p->loud = dalloc<int>((size_t)p->max_items, ok);
        p->quiet = dalloc<int>((size_t)p->max_items, ok);
        p->loud_cnt = dalloc<int>(4, ok);
        p->route = dalloc<int>((size_t)p->B, ok);
        p->split_cnt = dalloc<int>((size_t)(p->max_items / 256 + 1), ok);
        if (!ok) {
            for (void* b : {p->CfF, (void*)p->loud, (void*)p->quiet, (void*)p->loud_cnt, (void*)p->route,
                            (void*)p->split_cnt})
                if (b) (void)hipFree(b);
            p->CfF = nullptr;
            p->loud = nullptr;
            p->quiet = nullptr;
            p->loud_cnt = nullptr;
            p->route = nullptr;
            p->split_cnt = nullptr;
            return ERR_ALLOC;
        }
    }
    p->precision = precision;
    return OK;
}

extern "C" int bbhx_hip_direct_plan_enable_grad(bbhx_direct_plan* p)
{
    if (!p) return ERR_ARGS;
    if (p->Yadj) return OK;
    bool ok = true;
    p->Yadj = dalloc<double>((size_t)p->S * p->M * NA, ok);
    p->grecs = dalloc<gradk::GradRec>((size_t)p->B * p->M, ok);
    p->gpart = dalloc<double>((size_t)p->B * p->M * gradk::NGRAD, ok);
    if (!ok) {
        for (void* b : {(void*)p->Yadj, (void*)p->grecs, (void*)p->gpart})
            if (b) (void)hipFree(b);
        p->Yadj = nullptr;
        p->grecs = nullptr;
        p->gpart = nullptr;
        return ERR_ALLOC;
    }
    return OK;
}

extern "C" int bbhx_hip_direct_fused_grad(bbhx_direct_plan* p, double* d_h, double* h_h, double* grad, int* flags,
                                          const double* params, const double* t_start, const double* t_end,
                                          const int* ells_host, const int* mms_host, int numModes,
                                          const double* dataFreqs, const double* dataChannels,
                                          const double* noiseFactors, const double* qnm, int table_len, double dspin,
                                          double initial_t_val, double mf_min, double mf_max,
                                          const bbhx_response* orbits, int TDItag, int rescaled, int tdi2, int chunk,
                                          int block, int grid, void* stream)
{
    if (!grad) return ERR_ARGS;
    return direct_launch(p, d_h, h_h, grad, flags, params, t_start, t_end, ells_host, mms_host, numModes, dataFreqs,
                         dataChannels, noiseFactors, qnm, table_len, dspin, initial_t_val, mf_min, mf_max, orbits,
                         TDItag, rescaled, tdi2, chunk, block, grid, stream);
}
