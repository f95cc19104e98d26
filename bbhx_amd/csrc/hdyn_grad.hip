// Gradient of the heterodyned log-likelihood with respect to the 11 source
// parameters (ln M, q, chi1, chi2, d, phi_ref, cos inc, lam, sin beta, psi, t_ref).
//
// There is no counterpart in the reference (it has no derivatives). The value
// differentiated is the one HeterodynedLikelihood.get_ll returns
// (likelihood.py:550-648): logL = Re d_h - 1/2 Re h_h - 1/2 d_d.
//
// k_grad_setup_fwd the double records of the forward pass (as k_fused_setup)
//                and the per-node weights of the sums;
// k_grad_setup   (on a side stream, concurrently with the two kernels that
//                follow) the records with tangents in three independent parts,
//                each in wavefronts of its own: a PhenomHM record whose every
//                coefficient carries its derivatives with respect to
//                (ln M, q, chi1, chi2) (Dual<4> instantiation of the set-up
//                code; its mode-dependent part one lane per (template, mode),
//                its (2,2) reference part one lane per template), and the
//                polarisation basis, wave vector and Y factors with the frame's
//                rotation coefficients in lam, beta and the Y factors'
//                derivatives in cos inc.
// k_grad_weights one workgroup per template (as the forward kernel's):
//   1. forward: residuals r = h / h0 on the K nodes (LDS) and d_h, h_h,
//      exactly as the forward kernel;
//   2. adjoint of the heterodyned sums: for each node and channel the complex
//      weight g with  d logL = sum Re(g dh)  (from the per-node weights of
//      the segment formulas of Likelihood.cu:91-103 and r = h / h0), written
//      out (48 K bytes);
// k_grad_points  one 64-lane wavefront per (template, mode), lane per node:
//   3. per (node, mode) tangents: PhenomHM amplitude, phase and t(f) with
//      respect to the intrinsic parameters (structured tangents of the same
//      phm_eval code), the response by an adjoint of the (linear) TDI
//      combination and derivatives of the six single-link terms G_ij in t
//      and, through the frame's rotation, in lam, beta and psi, and in cos
//      inc through the Y factors; contracted with g at once, so each lane
//      keeps 11 accumulators; the wavefront reduces them;
// k_grad_reduce  the modes' partial gradients, summed in mode order.
// Splitting the forward pass from the per-point tangents keeps each kernel's
// register footprint (and so its occupancy) to its own needs.
// The spline and response are recomputed, nothing per node is stored.

#include "bbhx_hip.h"
#include "common.h"
#include "dual.h"
#include "fused_device.h"
#include "grad_device.h"
#include "workspace.h"

using namespace bbhx;
using namespace bbhx::fused;
using namespace bbhx::gradk;

namespace bbhx {
bool response_orbit_tab(const bbhx_response* r, resp::OrbitTab& o);
}

namespace {

// the double records of the forward pass, one lane per (template, mode), then
// the node weights (as k_fused_setup)
__global__ __launch_bounds__(64) void k_grad_setup_fwd(FusedRec* __restrict__ recs, NodeW* __restrict__ nw,
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

// the records with tangents over setup_blocks(B, M) blocks of 64 lanes, laid
// out as setup_split's: PART 0 and PART 2 one lane per (template, mode), PART 1
// (the (2,2) reference, which depends on the binary alone) one lane per
// template, written into each of its M records
__global__ __launch_bounds__(64) void k_grad_setup(GradRec* __restrict__ grecs, const double* __restrict__ params,
                                                   int B, ModeList modes, int M, const double* __restrict__ qnm,
                                                   int table_len, double dspin)
{
    const int nb = (B * M + 63) / 64, nt = (B + 63) / 64;
    const int blk = blockIdx.x;
    if (blk >= nb && blk < nb + nt) {
        const int b = (blk - nb) * 64 + threadIdx.x;
        if (b >= B) return;
        GradRec t;
        grad_setup_part<1>(t, params, B, b, 2, 2, qnm, table_len, dspin);
        for (int k = 0; k < M; k++) phm::copy_reference(grecs[(long)b * M + k].w, t.w);
        return;
    }
    const int part = blk < nb ? 0 : 2;
    const int idx = (part == 0 ? blk : blk - nb - nt) * 64 + threadIdx.x;
    if (idx >= B * M) return;
    const int b = idx / M, mi = idx % M;
    const int l = modes.ell[mi], m = modes.mm[mi];
    if (part == 0) grad_setup_part<0>(grecs[idx], params, B, b, l, m, qnm, table_len, dspin);
    else grad_setup_part<2>(grecs[idx], params, B, b, l, m, qnm, table_len, dspin);
}

// Adjoint weights of the heterodyned sums: for node n and channel c,
// g[c][n] = gamma / conj(h0) with d logL = sum_n,c Re(conj(gamma) dr) and
// r = h / h0, so that d logL = sum Re(g dh) for h = sum_m conj(X_m). In the
// per-node form of the sums (NodeW), logL = Re d_h - Re h_h / 2 - d_d / 2 with
// d_h = sum conj(r_n) W_n and Re h_h = sum |r_n|^2 Re V_n +
// Re(r_n conj(r_{n-1})) Re U_n, so
//   gamma_n = W_n - Re V_n r_n - (Re U_n r_{n-1} + Re U_{n+1} r_{n+1}) / 2
// (U_n = 0 where segment n is not summed, so node 0 has no r_{-1} term), and
// 1 / conj(h0) = conj(q).
__device__ inline void adjoint_weights(cd* gs, const cd* rs, const NodeW* __restrict__ nw, int K)
{
    for (int j = threadIdx.x; j < K; j += blockDim.x) {
        for (int c = 0; c < 3; c++) {
            const NodeW& w = nw[c * K + j];
            const cd r = rs[c * K + j];
            cd gam = w.w - w.v.re * r;
            if (j > 0) gam = gam - (0.5 * w.u.re) * rs[c * K + j - 1];
            if (j + 1 < K) gam = gam - (0.5 * nw[c * K + j + 1].u.re) * rs[c * K + j + 1];
            gs[c * K + j] = gam * conj(w.q);
        }
    }
}

// derivative of sinc at a (sinc' = (cos a - sinc a) / a; 0 at a = 0)
// Gradient contributions of one (node, mode) to acc[11]. g: adjoint weights of
// the node's three channels (d logL = sum_c Re(g_c dh_c), h_c = sum_m conj(X_mc),
// X = T A e^{i phi}), so that W_c = g_c A e^{i phi}, a_A = Re S, a_phi = -A Im S
// with S = sum_c g_c T_c e^{i phi}.
__device__ inline void grad_point(const GradRec& G, const resp::OrbitTab& o, double f, const Window& win,
                                  const cd g[3], int TDItag, int rescaled, int tdi2, double dist, double acc[NGRAD],
                                  const double2* tab)
{
    auto weights = [&](const cd T[3], cd E1, double Av, cd W[3], double& aA, double& aPh) {
        cd S1 = mk(0.0, 0.0);
        for (int c = 0; c < 3; c++) {
            S1 += g[c] * T[c] * E1;
            W[c] = (g[c] * E1) * Av;
        }
        aA = S1.re;
        aPh = -Av * S1.im;
    };
    grad_point_core(G, o, f, win, true, weights, TDItag, rescaled, tdi2, dist, acc, tab);
}

// Forward pass and adjoint weights of one template (one wavefront): d_h, h_h
// as the forward kernel computes them, and gw[(b * 3 + c) * K + j] = the
// complex weight g of node j, channel c.
__global__ __launch_bounds__(64 * MAX_RESIDUAL_WAVES) __attribute__((amdgpu_waves_per_eu(2))) void k_grad_weights(
    double* __restrict__ d_h, double* __restrict__ h_h, double* __restrict__ ll, double d_d, cd* __restrict__ gw,
    const FusedRec* __restrict__ recs, const double* __restrict__ params, const double* __restrict__ t_start,
    const double* __restrict__ t_end, int B, ModeList ml, int M, const double* __restrict__ freqs,
    const NodeW* __restrict__ nw, int K, resp::OrbitTab orb, int TDItag, int rescaled, int tdi2,
    double initial_t_val)
{
    extern __shared__ cd rs[];  // residuals [3][K]
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
    adjoint_weights(gw + (long)b * 3 * K, rs, nw, K);
    if (threadIdx.x == 0) {
        d_h[2 * b] = sums[0];
        d_h[2 * b + 1] = sums[1];
        h_h[2 * b] = sums[2];
        h_h[2 * b + 1] = sums[3];
        // log-likelihood as the forward kernel writes it (likelihood.py:640-648)
        if (ll) ll[b] = -0.5 * ((d_d + sums[2]) - 2.0 * sums[0]);
    }
}

// Gradient contributions of one (template, mode): one wavefront, lane l owns
// nodes l, l+64, ... (the record is the same for the whole wavefront);
// partial[(b * M + m) * 11 + p].
__global__ __launch_bounds__(64) void k_grad_points(double* __restrict__ partial, const GradRec* __restrict__ grecs,
                                                    const cd* __restrict__ gw, const double* __restrict__ params,
                                                    const double* __restrict__ t_start,
                                                    const double* __restrict__ t_end, int B, int M,
                                                    const double* __restrict__ freqs, int K, resp::OrbitTab orb,
                                                    int TDItag, int rescaled, int tdi2, double initial_t_val)
{
    const int bm = blockIdx.x;
    const int b = bm / M;
    Window win;
    win.t_ref = params[P_TREF * B + b];
    win.ts = t_start[b];
    win.te = t_end[b];
    win.dt_shift = win.t_ref - initial_t_val;
    const double dist = params[P_DIST * B + b];
    const GradRec& G = grecs[bm];
    const cd* g_b = gw + (long)b * 3 * K;
    __shared__ double2 tab[129];
    etab_to_lds(tab);
    __syncthreads();
    double acc[NGRAD];
    for (int p = 0; p < NGRAD; p++) acc[p] = 0.0;
    for (int j = threadIdx.x; j < K; j += 64) {
        const cd g[3] = {g_b[j], g_b[K + j], g_b[2 * K + j]};
        grad_point(G, orb, freqs[j], win, g, TDItag, rescaled, tdi2, dist, acc, tab);
    }
    for (int p = 0; p < NGRAD; p++) acc[p] = wsum(acc[p]);
    if (threadIdx.x == 0)
        for (int p = 0; p < NGRAD; p++) partial[(long)bm * NGRAD + p] = acc[p];
}

// grad[p * B + b] = sum over modes, in mode order
__global__ void k_grad_reduce(double* __restrict__ grad, const double* __restrict__ partial, int B, int M)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * NGRAD) return;
    const int p = t / B, b = t % B;
    double s = 0.0;
    for (int m = 0; m < M; m++) s += partial[((long)b * M + m) * NGRAD + p];
    grad[(long)p * B + b] = s;
}

// One side stream (default priority) and a fork / join event pair per device,
// created on first use and kept for the life of the process.
struct SideStream {
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

SideStream* side_stream()
{
    static SideStream per_dev[64];
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    SideStream& s = per_dev[dev];
    if (!s.stream) {
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess) return nullptr;
        if (hipEventCreateWithFlags(&s.fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s.join, hipEventDisableTiming) != hipSuccess)
            return nullptr;
    }
    return &s;
}

}  // namespace

extern "C" int bbhx_hip_hdyn_fused_grad(double* d_h, double* h_h, double* grad, const double* params,
                                        const double* t_start, const double* t_end, int numBinAll,
                                        const int* ells_host, const int* mms_host, int numModes, const double* freqs,
                                        const double* h0, const double* dataConstants, int K, const double* qnm,
                                        int table_len, double dspin, double initial_t_val,
                                        const bbhx_response* orbits, int TDItag, int rescaled, int tdi2,
                                        int include_last_segment, double* logL, double d_d, void* stream)
{
    resp::OrbitTab orb;
    if (!response_orbit_tab(orbits, orb)) return ERR_ARGS;
    if (numBinAll < 0 || K < 2 || K > 1024 || table_len < 2) return ERR_ARGS;
    ModeList ml;
    const int mst = check_modes(ells_host, mms_host, numModes, ml);
    if (mst != OK) return mst;
    if (numBinAll == 0) return OK;
    if (!d_h || !h_h || !grad || !params || !t_start || !t_end || !freqs || !h0 || !dataConstants || !qnm)
        return ERR_ARGS;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int nrec = numBinAll * numModes;
    // the double records, then the node weights (as the forward's workspace)
    char* fws = static_cast<char*>(
        workspace_get(WS_FUSED, sizeof(FusedRec) * (size_t)nrec + sizeof(NodeW) * 3 * (size_t)K, st));
    FusedRec* recs = reinterpret_cast<FusedRec*>(fws);
    NodeW* nw = reinterpret_cast<NodeW*>(fws + sizeof(FusedRec) * (size_t)nrec);
    GradRec* grecs = static_cast<GradRec*>(workspace_get(WS_GRAD, sizeof(GradRec) * (size_t)nrec, st));
    // adjoint weights (B, 3, K) complex, then partial gradients (B, M, 11)
    const size_t gw_bytes = sizeof(cd) * 3 * (size_t)K * numBinAll;
    char* scratch = static_cast<char*>(
        workspace_get(WS_GRAD_AUX, gw_bytes + sizeof(double) * NGRAD * (size_t)nrec, st));
    if (!fws || !grecs || !scratch) return ERR_ALLOC;
    cd* gw = reinterpret_cast<cd*>(scratch);
    double* partial = reinterpret_cast<double*>(scratch + gw_bytes);
    // The records with tangents are needed only by k_grad_points; their set-up
    // (few lanes, long serial chains) runs on a side stream, beside the forward
    // records and the weights pass on the caller's stream (event fork / join, so
    // the sequence stays capturable in a graph).
    SideStream* side = side_stream();
    if (!side) return ERR_LAUNCH;
    hipEvent_t ev[3];
    const bool timed = timer_take3(ev);
    if (timed) (void)hipEventRecord(ev[0], st);
    if (hipEventRecord(side->fork, st) != hipSuccess || hipStreamWaitEvent(side->stream, side->fork, 0) != hipSuccess)
        return ERR_LAUNCH;
    hipLaunchKernelGGL(k_grad_setup, dim3(setup_blocks(numBinAll, numModes)), dim3(64), 0, side->stream, grecs,
                       params,
                       numBinAll, ml, numModes, qnm, table_len, dspin);
    if (hipEventRecord(side->join, side->stream) != hipSuccess) return ERR_LAUNCH;
    const int jend = include_last_segment ? K : K - 1;
    hipLaunchKernelGGL(k_grad_setup_fwd, dim3(setup_blocks(numBinAll, numModes) + ceil_div(3 * K, 64)), dim3(64),
                       0, st, recs, nw, params, numBinAll, ml, numModes, qnm, table_len, dspin, freqs,
                       reinterpret_cast<const cd*>(h0), reinterpret_cast<const cd*>(dataConstants), K, jend);
    if (hipGetLastError() != hipSuccess) return ERR_LAUNCH;
    if (timed) (void)hipEventRecord(ev[1], st);
    hipLaunchKernelGGL(k_grad_weights, dim3(numBinAll), dim3(64 * residual_waves(numBinAll, K)),
                       sizeof(cd) * 3 * (size_t)K, st, d_h, h_h, logL, d_d, gw, recs, params, t_start, t_end,
                       numBinAll, ml, numModes, freqs, nw, K, orb, TDItag, rescaled, tdi2, initial_t_val);
    if (hipStreamWaitEvent(st, side->join, 0) != hipSuccess) return ERR_LAUNCH;
    hipLaunchKernelGGL(k_grad_points, dim3(nrec), dim3(64), 0, st, partial, grecs, gw, params, t_start, t_end,
                       numBinAll, numModes, freqs, K, orb, TDItag, rescaled, tdi2, initial_t_val);
    hipLaunchKernelGGL(k_grad_reduce, dim3(ceil_div((long)numBinAll * NGRAD, 256)), dim3(256), 0, st, grad, partial,
                       numBinAll, numModes);
    if (timed) (void)hipEventRecord(ev[2], st);
    return hipGetLastError() == hipSuccess ? OK : ERR_LAUNCH;
}
