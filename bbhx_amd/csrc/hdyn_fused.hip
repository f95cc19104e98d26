// Fused heterodyned log-likelihood on gfx950.
//
// The reference evaluates HeterodynedLikelihood.get_ll (likelihood.py:550-648)
// as a chain of synchronous steps over a 9 x B x M x K FP64 buffer in HBM:
// PhenomHM kernel -> host phase/t shift -> response kernel -> direct_sum kernel
// -> h / h0 on the host side -> hdyn kernel (one lane per template looping over
// all 3K nodes). Here the whole chain is two launches and only d_h, h_h leave
// the chip:
//
//   k_fused_setup  one lane per (template, mode): QNM ringdown frequencies,
//                  PhenomHM coefficient record, response H matrix
//   k_fused_eval   one 64-lane wavefront per template; lane l owns sparse nodes
//                  j = l, l+64, ...; the mode loop accumulates the three TDI
//                  channels in registers; r = h/h0; segment j needs r_{j-1},
//                  taken from the neighbouring lane by a cross-lane shuffle;
//                  the heterodyned sums are reduced across the wavefront.
//
// FP64 throughout ("plain" form). No atomics; the reduction order is fixed.

#include <vector>

#include "bbhx_hip.h"
#include "common.h"
#include "phenomhm_device.h"
#include "response_device.h"
#include "workspace.h"

using namespace bbhx;

namespace bbhx {
bool response_orbit_tab(const bbhx_response* r, resp::OrbitTab& o);
}

namespace {

constexpr int NPARAM = 11;
enum { P_M1, P_M2, P_CHI1, P_CHI2, P_DIST, P_PHI, P_INC, P_LAM, P_BETA, P_PSI, P_TREF };

struct FusedRec {
    phm::PhmRec w;
    resp::RespRec s;
};

struct ModeList {
    int ell[MAX_MODES];
    int mm[MAX_MODES];
};

__global__ void k_fused_setup(FusedRec* __restrict__ recs, const double* __restrict__ params, int B, ModeList modes,
                              int M, const double* __restrict__ qnm, int table_len, double dspin)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * M) return;
    const int b = idx / M, mi = idx % M;
    const double m1 = params[P_M1 * B + b], m2 = params[P_M2 * B + b];
    const double c1 = params[P_CHI1 * B + b], c2 = params[P_CHI2 * B + b];
    double finmass, finspin;
    phm::final_state(m1, m2, c1, c2, finmass, finspin);
    // PhenomD (2,2) ringdown from the QNM table (PhenomHMWaveform.cu:1963-1992)
    int k = int((finspin - (-1.0)) / dspin);
    k = k < 0 ? 0 : (k > table_len - 1 ? table_len - 1 : k);
    const double x = finspin - (dspin * k + (-1.0));
    const double x2 = x * x, x3 = x2 * x;
    const double* T = qnm;
    const double rdD = (T[k] + T[table_len + k] * x + T[2 * table_len + k] * x2 + T[3 * table_len + k] * x3) / finmass;
    const double dmD =
        (T[4 * table_len + k] + T[5 * table_len + k] * x + T[6 * table_len + k] * x2 + T[7 * table_len + k] * x3) /
        finmass;
    double rd, dm;
    phm::hm_ringdown(modes.ell[mi], modes.mm[mi], finmass, finspin, rd, dm);
    FusedRec r;
    phm::phm_setup(r.w, m1 * MSUN_SI, m2 * MSUN_SI, c1, c2, params[P_DIST * B + b], modes.ell[mi], modes.mm[mi], rd,
                   dm, rdD, dmD, false);
    resp::resp_setup(r.s, modes.ell[mi], modes.mm[mi], params[P_PHI * B + b], params[P_INC * B + b],
                     params[P_LAM * B + b], params[P_BETA * B + b], params[P_PSI * B + b]);
    recs[idx] = r;
}

__device__ inline double wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One 64-lane wavefront (one workgroup) per template. Dynamic LDS holds the
// residuals r[3][K] of this template.
__global__ __launch_bounds__(64) void k_fused_eval(double* __restrict__ d_h, double* __restrict__ h_h,
                                                   const FusedRec* __restrict__ recs,
                                                   const double* __restrict__ params,
                                                   const double* __restrict__ t_start,
                                                   const double* __restrict__ t_end, int B, int M,
                                                   const double* __restrict__ freqs, const cd* __restrict__ h0,
                                                   const cd* __restrict__ dc, int K, resp::OrbitTab orb, int TDItag,
                                                   int rescaled, int tdi2, int jend, double initial_t_val)
{
    extern __shared__ cd rs[];  // [3][K]
    const int lane = threadIdx.x;
    const int b = blockIdx.x;
    const double t_ref = params[P_TREF * B + b];
    const double ts = t_start[b], te = t_end[b];
    const double dt_shift = t_ref - initial_t_val;
    const FusedRec* rb = recs + (long)b * M;

    for (int j = lane; j < K; j += 64) {
        const double f = freqs[j];
        cd h[3] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
        for (int mi = 0; mi < M; mi++) {
            const FusedRec& r = rb[mi];
            double amp, phase, tf;
            phm::phm_eval(r.w, f, amp, phase, tf);
            // phenomhm.py:487-507: phase += 2 pi f (t_ref - t0); tf += t_ref
            phase = phase + ((f * dt_shift) * 2) * PI;
            tf = tf + t_ref;
            cd T[3];
            double dphi;
            resp::resp_eval(r.s, orb, f, tf, TDItag, rescaled, tdi2, T, dphi);
            phase = phase + dphi;
            if ((tf >= ts) && ((tf <= te) || (te <= 0.0))) {
                const cd ap = amp * cexpi_red(phase);
                h[0] += conj(T[0] * ap);
                h[1] += conj(T[1] * ap);
                h[2] += conj(T[2] * ap);
            }
        }
        // r = h / h0 (likelihood.py:603)
        for (int c = 0; c < 3; c++) {
            const cd d = h0[c * K + j];
            const double iden = rcp_pos(d.re * d.re + d.im * d.im);
            rs[c * K + j] = mk((h[c].re * d.re + h[c].im * d.im) * iden, (h[c].im * d.re - h[c].re * d.im) * iden);
        }
    }
    __syncthreads();

    // heterodyned sums (Likelihood.cu:52-111); segment j joins nodes j-1 and j
    double l1r = 0.0, l1i = 0.0, l2r = 0.0, l2i = 0.0;
    for (int j = lane + 1; j < jend; j += 64) {
        const double fj = freqs[j], fprev = freqs[j - 1];
        if (fprev == 0.0) continue;
        const double inv = 1.0 / (fj - fprev);
        const double fm = (fj + fprev) / 2.0;
        for (int c = 0; c < 3; c++) {
            const cd cur = rs[c * K + j], prev = rs[c * K + j - 1];
            const cd A0 = dc[(0 * 3 + c) * K + j], A1 = dc[(1 * 3 + c) * K + j];
            const cd B0 = dc[(2 * 3 + c) * K + j], B1 = dc[(3 * 3 + c) * K + j];
            const cd r1 = mk((cur.re - prev.re) * inv, (cur.im - prev.im) * inv);
            const cd r0 = cur - r1 * (fj - fm);
            const cd r1c = conj(r1);
            const cd t1 = A0 * conj(r0) + A1 * r1c;
            const double mag2 = r0.re * r0.re + r0.im * r0.im;
            const cd t2 = B0 * mag2 + 2. * B1 * (r0 * r1c).re;
            l1r += t1.re;
            l1i += t1.im;
            l2r += t2.re;
            l2i += t2.im;
        }
    }
    l1r = wsum(l1r);
    l1i = wsum(l1i);
    l2r = wsum(l2r);
    l2i = wsum(l2i);
    if (lane == 0) {
        d_h[2 * b] = l1r;
        d_h[2 * b + 1] = l1i;
        h_h[2 * b] = l2r;
        h_h[2 * b + 1] = l2i;
    }
}

// Optional event timing of the fused launches (bbhx_hip_timing_*): three events per
// call (before setup, between, after eval), read back only on request.
struct FusedTimer {
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    hipEvent_t take()
    {
        if (used == pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[used++];
    }
};
FusedTimer g_timer;

}  // namespace

extern "C" int bbhx_hip_timing_enable(int on)
{
    g_timer.on = on != 0;
    g_timer.used = 0;
    return OK;
}

extern "C" int bbhx_hip_timing_read(double* ms_setup, double* ms_eval, int* ncalls)
{
    double s = 0.0, e = 0.0;
    int n = 0;
    for (size_t k = 0; k + 2 < g_timer.used; k += 3) {
        float a = 0.f, b = 0.f;
        if (hipEventSynchronize(g_timer.pool[k + 2]) != hipSuccess) return ERR_LAUNCH;
        if (hipEventElapsedTime(&a, g_timer.pool[k], g_timer.pool[k + 1]) != hipSuccess) return ERR_LAUNCH;
        if (hipEventElapsedTime(&b, g_timer.pool[k + 1], g_timer.pool[k + 2]) != hipSuccess) return ERR_LAUNCH;
        s += a;
        e += b;
        n++;
    }
    if (ms_setup) *ms_setup = s;
    if (ms_eval) *ms_eval = e;
    if (ncalls) *ncalls = n;
    g_timer.used = 0;
    return OK;
}

extern "C" int bbhx_hip_hdyn_fused(double* d_h, double* h_h, const double* params, const double* t_start,
                                   const double* t_end, int numBinAll, const int* ells_host, const int* mms_host,
                                   int numModes, const double* freqs, const double* h0, const double* dataConstants,
                                   int K, const double* qnm, int table_len, double dspin, double initial_t_val,
                                   const bbhx_response* orbits, int TDItag, int rescaled, int tdi2,
                                   int include_last_segment, int precision, void* stream)
{
    resp::OrbitTab orb;
    if (!response_orbit_tab(orbits, orb)) return ERR_ARGS;
    if (numBinAll < 0 || numModes < 1 || numModes > MAX_MODES || K < 2 || K > 1024 || table_len < 2) return ERR_ARGS;
    if (precision != 0) return ERR_ARGS;
    if (numBinAll == 0) return OK;
    if (!d_h || !h_h || !params || !t_start || !t_end || !ells_host || !mms_host || !freqs || !h0 ||
        !dataConstants || !qnm)
        return ERR_ARGS;
    ModeList ml{};
    for (int m = 0; m < numModes; m++) {
        const int l = ells_host[m], mm = mms_host[m];
        if (!((l == 2 && mm == 2) || (l == 2 && mm == 1) || (l == 3 && mm == 3) || (l == 3 && mm == 2) ||
              (l == 4 && mm == 4) || (l == 4 && mm == 3)))
            return ERR_MODE;
        ml.ell[m] = l;
        ml.mm[m] = mm;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int nrec = numBinAll * numModes;
    FusedRec* recs = static_cast<FusedRec*>(workspace_get(WS_FUSED, sizeof(FusedRec) * (size_t)nrec));
    if (!recs) return ERR_ALLOC;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    if (g_timer.on) {
        ev0 = g_timer.take();
        ev1 = g_timer.take();
        ev2 = g_timer.take();
        if (!ev0 || !ev1 || !ev2) return ERR_LAUNCH;
        (void)hipEventRecord(ev0, st);
    }
    hipLaunchKernelGGL(k_fused_setup, dim3(ceil_div(nrec, 64)), dim3(64), 0, st, recs, params, numBinAll, ml,
                       numModes, qnm, table_len, dspin);
    if (hipGetLastError() != hipSuccess) return ERR_LAUNCH;
    if (g_timer.on) (void)hipEventRecord(ev1, st);
    const int jend = include_last_segment ? K : K - 1;
    const size_t lds = sizeof(cd) * 3 * (size_t)K;
    hipLaunchKernelGGL(k_fused_eval, dim3(numBinAll), dim3(64), lds, st, d_h, h_h, recs, params, t_start, t_end,
                       numBinAll, numModes, freqs, reinterpret_cast<const cd*>(h0),
                       reinterpret_cast<const cd*>(dataConstants), K, orb, TDItag, rescaled, tdi2, jend,
                       initial_t_val);
    if (g_timer.on) (void)hipEventRecord(ev2, st);
    return hipGetLastError() == hipSuccess ? OK : ERR_LAUNCH;
}

extern "C" int bbhx_hip_release_workspace(void)
{
    workspace_release_all();
    return OK;
}

extern "C" const char* bbhx_hip_version(void) { return "bbhx-hip 0.1 (gfx950)"; }
