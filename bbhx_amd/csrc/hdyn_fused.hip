// Fused heterodyned log-likelihood on gfx950.
//
// The reference evaluates HeterodynedLikelihood.get_ll (likelihood.py:550-648)
// as a chain of synchronous steps over a 9 x B x M x K FP64 buffer in HBM:
// PhenomHM kernel -> host phase/t shift -> response kernel -> direct_sum kernel
// -> h / h0 on the host side -> hdyn kernel (one lane per template looping over
// all 3K nodes). Here the whole chain is two launches and only d_h, h_h leave
// the chip:
//
//   k_fused_setup  the (template, mode) records in three independent parts
//                  (mode-dependent PhenomHM coefficients and QNM ringdown, one
//                  lane per record; the (2,2) reference, one lane per
//                  template; the response basis and Y factors, one lane per
//                  record), and in extra blocks the per-node weights of the
//                  heterodyned sums and 1 / h0 (NodeW)
//   k_fused_eval   one workgroup per template (one 64-lane wavefront at large
//                  batches); lane l owns sparse nodes j = l, l + 64, ...; the
//                  records are copied into LDS; per node the modes' link terms
//                  are summed before one TDI combination; r = h / h0 goes to
//                  LDS; after one barrier each lane adds its nodes' terms of
//                  the heterodyned sums (per-node form) and the sums are
//                  reduced across the workgroup in a fixed order.
//
// FP64 throughout ("plain" form). No atomics; the reduction order is fixed.

#include <map>
#include <mutex>
#include <vector>

#include "bbhx_hip.h"
#include "common.h"
#include "fused_device.h"
#include "workspace.h"

using namespace bbhx;
using namespace bbhx::fused;

namespace bbhx {
bool response_orbit_tab(const bbhx_response* r, resp::OrbitTab& o);
}

namespace {

// The records over setup_blocks(B, M) blocks, then the node weights
// over ceil(3 K / 64) more.
__global__ __launch_bounds__(64) void k_fused_setup(FusedRec* __restrict__ recs, NodeW* __restrict__ nw,
                                                    const double* __restrict__ params, int B, ModeList modes, int M,
                                                    const double* __restrict__ qnm, int table_len, double dspin,
                                                    const double* __restrict__ freqs, const cd* __restrict__ h0,
                                                    const cd* __restrict__ dc, int K, int jend)
{
    const int nsb = setup_blocks(B, M);
    if ((int)blockIdx.x >= nsb) {
        node_weight(nw, freqs, h0, dc, K, jend, ((int)blockIdx.x - nsb) * 64 + threadIdx.x);
        return;
    }
    setup_split(recs, params, B, modes, M, qnm, table_len, dspin);
}

// One workgroup per template, of residual_waves(B, K) wavefronts (one at large
// batches; several at small ones, so that the chip still holds two per SIMD).
// Dynamic LDS holds the residuals r[3][K] of this template; static LDS its
// records and the e^{i theta} table (the records read with LDS broadcast loads:
// measured faster than scalar loads, whose waits the SGPR budget serialises).
// The register budget is capped so that two wavefronts share each SIMD (256
// VGPRs each, no spills); capped at three per SIMD (168 VGPRs) values spill and
// it takes twice as long.
__global__ __launch_bounds__(64 * MAX_RESIDUAL_WAVES) __attribute__((amdgpu_waves_per_eu(2))) void k_fused_eval(
    double* __restrict__ d_h, double* __restrict__ h_h, double* __restrict__ ll, double d_d,
    const FusedRec* __restrict__ recs, const double* __restrict__ params, const double* __restrict__ t_start,
    const double* __restrict__ t_end, int B, ModeList ml, int M, const double* __restrict__ freqs,
    const NodeW* __restrict__ nw, int K, resp::OrbitTab orb, int TDItag, int rescaled, int tdi2,
    double initial_t_val)
{
    extern __shared__ cd rs[];  // [3][K]
    __shared__ double2 tab[129];
    etab_to_lds(tab);
    const int b = blockIdx.x;
    Window win;
    win.t_ref = params[P_TREF * B + b];
    win.ts = t_start[b];
    win.te = t_end[b];
    win.dt_shift = win.t_ref - initial_t_val;
    __shared__ FusedRec srec[MAX_MODES];
    records_to_lds(srec, recs + (long)b * M, M);
    __syncthreads();
    residuals(rs, srec, ml, M, win, freqs, nw, K, orb, TDItag, rescaled, tdi2, tab);
    __syncthreads();
    double sums[4];
    node_sums(rs, nw, K, sums);
    if (threadIdx.x == 0) {
        d_h[2 * b] = sums[0];
        d_h[2 * b + 1] = sums[1];
        h_h[2 * b] = sums[2];
        h_h[2 * b + 1] = sums[3];
        // log-likelihood (likelihood.py:640-648): 2 Re d_h is exact, so this is
        // the value of the same expression formed from d_h, h_h afterwards
        if (ll) ll[b] = -0.5 * ((d_d + sums[2]) - 2.0 * sums[0]);
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

namespace bbhx {
namespace fused {

bool timer_take3(hipEvent_t ev[3])
{
    if (!g_timer.on) return false;
    for (int i = 0; i < 3; i++) {
        ev[i] = g_timer.take();
        if (!ev[i]) return false;
    }
    return true;
}

int check_modes(const int* ells_host, const int* mms_host, int numModes, ModeList& ml)
{
    if (!ells_host || !mms_host || numModes < 1 || numModes > MAX_MODES) return ERR_ARGS;
    ml = ModeList{};
    for (int m = 0; m < numModes; m++) {
        const int l = ells_host[m], mm = mms_host[m];
        if (!((l == 2 && mm == 2) || (l == 2 && mm == 1) || (l == 3 && mm == 3) || (l == 3 && mm == 2) ||
              (l == 4 && mm == 4) || (l == 4 && mm == 3)))
            return ERR_MODE;
        ml.ell[m] = l;
        ml.mm[m] = mm;
    }
    // evaluation order: each mode followed by the next unplaced mode of equal m
    bool placed[MAX_MODES] = {};
    int k = 0;
    for (int m = 0; m < numModes; m++) {
        if (placed[m]) continue;
        placed[m] = true;
        ml.order[k] = m;
        ml.pair[k++] = 0;
        for (int n = m + 1; n < numModes; n++)
            if (!placed[n] && ml.mm[n] == ml.mm[m]) {
                placed[n] = true;
                ml.pair[k - 1] = 1;
                ml.order[k] = n;
                ml.pair[k++] = 0;
                break;
            }
    }
    return OK;
}

}  // namespace fused
}  // namespace bbhx

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

namespace {

// The record set-up and evaluation launches of the fused forward, on records
// 'recs' (numBinAll * numModes of them) and node weights 'nw' (3 K) that the
// caller provides; arguments as bbhx_hip_hdyn_fused, already checked. 'timing'
// allows the optional event timing (which may create events).
int hdyn_fused_launch(FusedRec* recs, NodeW* nw, double* d_h, double* h_h, const double* params, const double* t_start,
                      const double* t_end, int numBinAll, const ModeList& ml, int numModes, const double* freqs,
                      const double* h0, const double* dataConstants, int K, const double* qnm, int table_len,
                      double dspin, double initial_t_val, const resp::OrbitTab& orb, int TDItag, int rescaled,
                      int tdi2, int include_last_segment, double* logL, double d_d, hipStream_t st, bool timing)
{
    const int nrec = numBinAll * numModes;
    hipEvent_t ev[3];
    const bool timed = timing && timer_take3(ev);
    if (timing && g_timer.on && !timed) return ERR_LAUNCH;
    if (timed) (void)hipEventRecord(ev[0], st);
    const int jend = include_last_segment ? K : K - 1;
    hipLaunchKernelGGL(k_fused_setup, dim3(setup_blocks(numBinAll, numModes) + ceil_div(3 * K, 64)), dim3(64), 0, st,
                       recs, nw, params, numBinAll, ml, numModes, qnm, table_len, dspin, freqs,
                       reinterpret_cast<const cd*>(h0), reinterpret_cast<const cd*>(dataConstants), K, jend);
    if (hipGetLastError() != hipSuccess) return ERR_LAUNCH;
    if (timed) (void)hipEventRecord(ev[1], st);
    const size_t lds = sizeof(cd) * 3 * (size_t)K;
    hipLaunchKernelGGL(k_fused_eval, dim3(numBinAll), dim3(64 * residual_waves(numBinAll, K)), lds, st, d_h, h_h,
                       logL, d_d, recs, params, t_start, t_end, numBinAll, ml, numModes, freqs, nw, K, orb, TDItag,
                       rescaled, tdi2, initial_t_val);
    if (timed) (void)hipEventRecord(ev[2], st);
    return hipGetLastError() == hipSuccess ? OK : ERR_LAUNCH;
}

// Argument checks shared by the fused forward entry points. Returns OK (and
// fills orb, ml), a status, or -1 for an empty batch (nothing to do).
int hdyn_fused_check(double* d_h, double* h_h, const double* params, const double* t_start, const double* t_end,
                     int numBinAll, const int* ells_host, const int* mms_host, int numModes, const double* freqs,
                     const double* h0, const double* dataConstants, int K, const double* qnm, int table_len,
                     const bbhx_response* orbits, int precision, resp::OrbitTab& orb, ModeList& ml)
{
    if (!response_orbit_tab(orbits, orb)) return ERR_ARGS;
    if (numBinAll < 0 || numModes < 1 || numModes > MAX_MODES || K < 2 || K > 1024 || table_len < 2) return ERR_ARGS;
    if (precision != 0) return ERR_ARGS;
    if (numBinAll == 0) return -1;
    if (!d_h || !h_h || !params || !t_start || !t_end || !ells_host || !mms_host || !freqs || !h0 ||
        !dataConstants || !qnm)
        return ERR_ARGS;
    return check_modes(ells_host, mms_host, numModes, ml);
}

// ---- the allocation-free form: records in a fixed device array of the library

constexpr long NOALLOC_MAX_RECORDS = BBHX_HIP_NOALLOC_MAX_RECORDS;

// One copy per device, held until the process ends. In host code the name is
// only a stand-in: its device address comes from hipGetSymbolAddress.
// The records, then the node weights for up to 1024 nodes.
constexpr size_t NOALLOC_NW_OFFSET = sizeof(FusedRec) * NOALLOC_MAX_RECORDS;
static __device__ __align__(256) unsigned char scratch[NOALLOC_NW_OFFSET + sizeof(NodeW) * 3 * 1024];

std::mutex g_scratch_mutex;
std::map<int, void*> g_scratch_addr;  // device -> address of 'scratch' on that device

// Address of 'scratch' on the current device, looked up once per device (with
// that device current) and kept; nullptr if the lookup fails.
void* scratch_address()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    auto it = g_scratch_addr.find(dev);
    if (it != g_scratch_addr.end()) return it->second;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        (void)hipGetLastError();
        return nullptr;
    }
    g_scratch_addr[dev] = p;
    return p;
}

}  // namespace

extern "C" int bbhx_hip_hdyn_fused(double* d_h, double* h_h, const double* params, const double* t_start,
                                   const double* t_end, int numBinAll, const int* ells_host, const int* mms_host,
                                   int numModes, const double* freqs, const double* h0, const double* dataConstants,
                                   int K, const double* qnm, int table_len, double dspin, double initial_t_val,
                                   const bbhx_response* orbits, int TDItag, int rescaled, int tdi2,
                                   int include_last_segment, int precision, double* logL, double d_d,
                                   void* stream)
{
    resp::OrbitTab orb;
    ModeList ml;
    const int chk = hdyn_fused_check(d_h, h_h, params, t_start, t_end, numBinAll, ells_host, mms_host, numModes,
                                     freqs, h0, dataConstants, K, qnm, table_len, orbits, precision, orb, ml);
    if (chk != OK) return chk < 0 ? OK : chk;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int nrec = numBinAll * numModes;
    // the records, then the node weights
    char* ws = static_cast<char*>(
        workspace_get(WS_FUSED, sizeof(FusedRec) * (size_t)nrec + sizeof(NodeW) * 3 * (size_t)K, st));
    if (!ws) return ERR_ALLOC;
    FusedRec* recs = reinterpret_cast<FusedRec*>(ws);
    NodeW* nw = reinterpret_cast<NodeW*>(ws + sizeof(FusedRec) * (size_t)nrec);
    return hdyn_fused_launch(recs, nw, d_h, h_h, params, t_start, t_end, numBinAll, ml, numModes, freqs, h0,
                             dataConstants, K, qnm, table_len, dspin, initial_t_val, orb, TDItag, rescaled, tdi2,
                             include_last_segment, logL, d_d, st, true);
}

extern "C" int bbhx_hip_hdyn_fused_noalloc(double* d_h, double* h_h, const double* params, const double* t_start,
                                           const double* t_end, int numBinAll, const int* ells_host,
                                           const int* mms_host, int numModes, const double* freqs, const double* h0,
                                           const double* dataConstants, int K, const double* qnm, int table_len,
                                           double dspin, double initial_t_val, const bbhx_response* orbits,
                                           int TDItag, int rescaled, int tdi2, int include_last_segment,
                                           int precision, double* logL, double d_d, void* stream)
{
    resp::OrbitTab orb;
    ModeList ml;
    const int chk = hdyn_fused_check(d_h, h_h, params, t_start, t_end, numBinAll, ells_host, mms_host, numModes,
                                     freqs, h0, dataConstants, K, qnm, table_len, orbits, precision, orb, ml);
    if (chk != OK) return chk < 0 ? OK : chk;
    if ((long)numBinAll * numModes > NOALLOC_MAX_RECORDS) return ERR_CAPACITY;
    char* base = static_cast<char*>(scratch_address());
    if (!base) return ERR_LAUNCH;
    FusedRec* recs = reinterpret_cast<FusedRec*>(base);
    NodeW* nw = reinterpret_cast<NodeW*>(base + NOALLOC_NW_OFFSET);
    return hdyn_fused_launch(recs, nw, d_h, h_h, params, t_start, t_end, numBinAll, ml, numModes, freqs, h0,
                             dataConstants, K, qnm, table_len, dspin, initial_t_val, orb, TDItag, rescaled, tdi2,
                             include_last_segment, logL, d_d, reinterpret_cast<hipStream_t>(stream), false);
}

namespace {

// Angular rate of the constellation's guiding centre in the SSB-frame time
// transform (ConstOmega, utils/constants.py:48, used by transform.py:75-105);
// not the orbit's 2 pi / year of the response.
constexpr double OMEGA_LFRAME = 1.99098659277e-7;

// SSB-frame observation window of each binary (waveformbuild.py:477-484):
// t = tL + R cos(beta) cos(p) - Omega/2 (R cos(beta))^2 sin(2 p), p = Omega tL - lam,
// R = 1 AU / c (OrbitR / C_SI, transform.py:25, :99; tSSBfromLframe with t0 = 0,
// transform.py:75-105), at tL = tL0 and tL1.
__global__ void k_window_times(double* __restrict__ ts, double* __restrict__ te, const double* __restrict__ lam,
                               const double* __restrict__ beta, int B, double tL0, double tL1)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const double rc = (AU_SI / C_SI) * cos(beta[b]);
    const double c2 = (0.5 * OMEGA_LFRAME) * (rc * rc);
    const double l = lam[b];
    const double p0 = OMEGA_LFRAME * tL0 - l, p1 = OMEGA_LFRAME * tL1 - l;
    ts[b] = (tL0 + rc * cos(p0)) - c2 * sin(2.0 * p0);
    te[b] = (tL1 + rc * cos(p1)) - c2 * sin(2.0 * p1);
}

}  // namespace

extern "C" int bbhx_hip_window_times(double* t_start, double* t_end, const double* lam, const double* beta,
                                     int numBinAll, double tL_start, double tL_end, void* stream)
{
    if (numBinAll < 0 || (numBinAll > 0 && (!t_start || !t_end || !lam || !beta))) return ERR_ARGS;
    if (numBinAll == 0) return OK;
    hipLaunchKernelGGL(k_window_times, dim3(ceil_div(numBinAll, 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), t_start, t_end, lam, beta, numBinAll, tL_start, tL_end);
    return hipGetLastError() == hipSuccess ? OK : ERR_LAUNCH;
}

extern "C" int bbhx_hip_release_workspace(void)
{
    workspace_release_all();
    return OK;
}

extern "C" const char* bbhx_hip_version(void) { return "bbhx-hip 0.1 (gfx950)"; }
