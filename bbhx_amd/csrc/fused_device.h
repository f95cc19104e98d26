// Device pieces shared by the fused heterodyned-likelihood kernels (forward
// and gradient): the per-(template, mode) record set-up, the residual pass
// r = h / h0 on the sparse nodes and the heterodyned segment sums.
//
// Reference chain: HeterodynedLikelihood.get_ll (likelihood.py:550-648) ->
// PhenomHM (PhenomHMWaveform.cu), t_ref shift (phenomhm.py:487-507), response
// (Response.cu), windowed mode sum (WaveformBuild.cu:48-60), r = h / h0
// (likelihood.py:603), hdyn (Likelihood.cu:52-111).
#pragma once

#include "common.h"
#include "dual.h"
#include "phenomhm_device.h"
#include "response_device.h"

namespace bbhx {
namespace fused {

constexpr int NPARAM = 11;
enum { P_M1, P_M2, P_CHI1, P_CHI2, P_DIST, P_PHI, P_INC, P_LAM, P_BETA, P_PSI, P_TREF };

struct FusedRec {
    phm::PhmRec w;
    resp::RespRec s;
};

// The modes, and the order in which the residual pass evaluates them: modes
// of equal m next to each other, pair[k] = 1 where order[k + 1] has the m of
// order[k] and is evaluated with it (see residuals).
struct ModeList {
    int ell[MAX_MODES];
    int mm[MAX_MODES];
    int order[MAX_MODES];
    int pair[MAX_MODES];
};

// PhenomD (2,2) ringdown frequencies in units of 1/M from the QNM table spline
// (PhenomHMWaveform.cu:1963-1992); the spin index is clamped into the table.
template <class S>
__device__ inline void phenomd_ringdown(S finmass, S finspin, const double* __restrict__ T, int table_len,
                                        double dspin, S& rd, S& dm)
{
    int k = int((value(finspin) - (-1.0)) / dspin);
    k = k < 0 ? 0 : (k > table_len - 1 ? table_len - 1 : k);
    const S x = finspin - (dspin * k + (-1.0));
    const S x2 = x * x, x3 = x2 * x;
    rd = (T[k] + T[table_len + k] * x + T[2 * table_len + k] * x2 + T[3 * table_len + k] * x3) / finmass;
    dm = (T[4 * table_len + k] + T[5 * table_len + k] * x + T[6 * table_len + k] * x2 + T[7 * table_len + k] * x3) /
         finmass;
}

// Record of template b (params: (11, B) rows), mode (l, m), in parts.
// One of the three independent dependency chains of a record's set-up (they
// write disjoint fields of the record): PART 0, the mode-dependent PhenomHM
// coefficients; PART 1, the (2,2) reference (PhenomD amplitude, Mf_ref, phi0,
// t0), the same for every mode of a binary; PART 2, the response record. A
// set-up kernel gives each part its own wavefronts (a branch inside one
// wavefront would run them one after another), so the set-up takes the longest
// chain instead of their sum.

template <int PART>
__device__ inline void setup_part(FusedRec& r, const double* __restrict__ params, int B, int b, int l, int m,
                                  const double* __restrict__ qnm, int table_len, double dspin)
{
    if constexpr (PART == 2) {
        resp::resp_setup(r.s, l, m, params[P_PHI * B + b], params[P_INC * B + b], params[P_LAM * B + b],
                         params[P_BETA * B + b], params[P_PSI * B + b]);
    } else {
        const double m1 = params[P_M1 * B + b], m2 = params[P_M2 * B + b];
        const double c1 = params[P_CHI1 * B + b], c2 = params[P_CHI2 * B + b];
        double finmass, finspin;
        phm::final_state(m1, m2, c1, c2, finmass, finspin);
        double rdD, dmD, rd = 0.0, dm = 0.0;
        phenomd_ringdown(finmass, finspin, qnm, table_len, dspin, rdD, dmD);
        if constexpr (PART == 0) phm::hm_ringdown(l, m, finmass, finspin, rd, dm);
        phm::phm_setup<double, PART>(r.w, m1 * MSUN_SI, m2 * MSUN_SI, c1, c2, params[P_DIST * B + b], l, m, rd,
                                     dm, rdD, dmD, false);
    }
}

// PART 1 (the (2,2) reference) depends on the binary alone: computed once per
// template and written into each of its M records.
__device__ inline void setup_reference(FusedRec* __restrict__ rb, int M, const double* __restrict__ params, int B,
                                       int b, const double* __restrict__ qnm, int table_len, double dspin)
{
    FusedRec t;
    setup_part<1>(t, params, B, b, 2, 2, qnm, table_len, dspin);
    for (int k = 0; k < M; k++) phm::copy_reference(rb[k].w, t.w);
}

// Blocks of 64 lanes of a set-up kernel for B templates of M modes: PART 0 and
// PART 2 one lane per record (nb = ceil(B M / 64) blocks each), PART 1 one lane
// per template (ceil(B / 64) blocks).
__host__ __device__ inline int setup_blocks(int B, int M) { return 2 * ((B * M + 63) / 64) + (B + 63) / 64; }

// Body of a set-up kernel of 64-thread workgroups over setup_blocks(B, M)
// blocks: [0, nb) PART 0, [nb, nb + nt) PART 1, then PART 2 (nb = ceil(B M /
// 64), nt = ceil(B / 64)).
__device__ inline void setup_split(FusedRec* __restrict__ recs, const double* __restrict__ params, int B,
                                   const ModeList& modes, int M, const double* __restrict__ qnm, int table_len,
                                   double dspin)
{
    const int nb = (B * M + 63) / 64, nt = (B + 63) / 64;
    const int blk = blockIdx.x;
    if (blk >= nb && blk < nb + nt) {
        const int b = (blk - nb) * 64 + threadIdx.x;
        if (b < B) setup_reference(recs + (long)b * M, M, params, B, b, qnm, table_len, dspin);
        return;
    }
    const int part = blk < nb ? 0 : 2;
    const int idx = (part == 0 ? blk : blk - nb - nt) * 64 + threadIdx.x;
    if (idx >= B * M) return;
    const int b = idx / M, mi = idx % M;
    const int l = modes.ell[mi], m = modes.mm[mi];
    if (part == 0) setup_part<0>(recs[idx], params, B, b, l, m, qnm, table_len, dspin);
    else setup_part<2>(recs[idx], params, B, b, l, m, qnm, table_len, dspin);
}

__device__ inline double wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

constexpr int MAX_RESIDUAL_WAVES = 8;

struct Window {
    double t_ref, ts, te, dt_shift;
};

__device__ inline bool in_window(const Window& w, double tf)
{
    return (tf >= w.ts) && ((tf <= w.te) || (w.te <= 0.0));
}

// The heterodyned sums (Likelihood.cu:91-103) per segment j of nodes j-1, j
// with r1 = (r_j - r_{j-1}) / (f_j - f_{j-1}) and r0 = r at the segment's
// midpoint, (r_j + r_{j-1}) / 2, are
//   d_h = sum_j A0 conj(r0) + A1 conj(r1),
//   h_h = sum_j B0 |r0|^2 + 2 B1 Re(r0 conj(r1)),
// and since Re(r0 conj(r1)) = (|r_j|^2 - |r_{j-1}|^2) / (2 (f_j - f_{j-1})) and
// |r0|^2 = (|r_j|^2 + |r_{j-1}|^2 + 2 Re(r_j conj(r_{j-1}))) / 4, per node n
//   d_h = sum_n conj(r_n) W_n,
//   h_h = sum_n |r_n|^2 V_n + Re(r_n conj(r_{n-1})) U_n
// with W_n = [A0/2 + A1 i]_n + [A0/2 - A1 i]_{n+1}, V_n = [B0/4 + B1 i]_n +
// [B0/4 - B1 i]_{n+1}, U_n = [B0/2]_n, i = 1 / (f_j - f_{j-1}), each bracket the
// segment's term (0 for a segment the sums skip). The weights depend on the
// nodes and data alone, so they are formed once per call; q = 1 / h0 of the
// node (r = h q). One entry per (channel c, node n) at c * K + n.
struct NodeW {
    cd w, v, u, q;
};

// Entry idx (< 3 K) of the node weights from the heterodyne constants
// dc[(q * 3 + c) * K + j] (A0, A1, B0, B1 of segment j) and the reference h0;
// segments j < jend whose lower node is not at f = 0 are summed.
__device__ inline void node_weight(NodeW* __restrict__ nw, const double* __restrict__ freqs,
                                   const cd* __restrict__ h0, const cd* __restrict__ dc, int K, int jend, int idx)
{
    if (idx >= 3 * K) return;
    const int c = idx / K, n = idx - c * K;
    auto valid = [&](int j) { return j >= 1 && j < jend && freqs[j - 1] != 0.0; };
    cd W = mk(0.0, 0.0), V = mk(0.0, 0.0), U = mk(0.0, 0.0);
    if (valid(n)) {
        const double inv = rcp_pos(freqs[n] - freqs[n - 1]);
        const cd A0 = dc[(0 * 3 + c) * K + n], A1 = dc[(1 * 3 + c) * K + n];
        const cd B0 = dc[(2 * 3 + c) * K + n], B1 = dc[(3 * 3 + c) * K + n];
        W = 0.5 * A0 + A1 * inv;
        V = 0.25 * B0 + B1 * inv;
        U = 0.5 * B0;
    }
    if (n + 1 < K && valid(n + 1)) {
        const double inv = rcp_pos(freqs[n + 1] - freqs[n]);
        const cd A0 = dc[(0 * 3 + c) * K + n + 1], A1 = dc[(1 * 3 + c) * K + n + 1];
        const cd B0 = dc[(2 * 3 + c) * K + n + 1], B1 = dc[(3 * 3 + c) * K + n + 1];
        W = W + (0.5 * A0 - A1 * inv);
        V = V + (0.25 * B0 - B1 * inv);
    }
    const cd d = h0[c * K + n];
    const double iden = rcp_pos(d.re * d.re + d.im * d.im);
    nw[idx] = NodeW{W, V, U, mk(d.re * iden, -d.im * iden)};
}

// Residual pass of one template by one workgroup: thread t owns nodes t,
// t + blockDim.x, ...; r = h / h0 is written to rs[c * K + j] (LDS).
// Per node the modes are summed before the TDI combination: with link terms
// G_m and weights w_m = A_m e^{i phi_m} (0 outside the observation window),
// sum_m conj(T_c(G_m) w_m) = conj(T_c(sum_m w_m G_m)) because the TDI
// combination T_c is linear in the link terms with coefficients that depend
// on the frequency alone (WaveformBuild.cu:48-60 sums conj(T A e^{i phi}) mode
// by mode; the same sum, reassociated). 'tab' is the e^{i theta} table.
//
// Two modes of equal m (the (2,2) and (3,2), the (3,3) and (4,3)) have the same
// PhenomD amplitude, phase and t(f) wherever both are on the first piece of
// their frequency maps (Mf = (2/m) f M for both) and the phase on its inspiral
// piece: those use only coefficients that do not depend on the mode's ringdown
// (the (2,2) amplitude, the TaylorF2 and inspiral fits, t0, phi0), so the
// second mode's amplitude, phase and t(f) equal the first's except for its PN
// amplitude factor, bit for bit as its own evaluation would give them. When a
// whole wavefront is there, the pair is evaluated once: one PhenomHM
// evaluation, the second mode's factor, and one response evaluation with the
// two modes' Y factors weighted by their amplitudes.
template <class E>
__device__ inline void residuals(cd* rs, const FusedRec* __restrict__ rb, const ModeList& ml, int M,
                                 const Window& win, const double* __restrict__ freqs, const NodeW* __restrict__ nw,
                                 int K, const resp::OrbitTab& orb, int TDItag, int rescaled, int tdi2,
                                 const E* __restrict__ tab)
{
    for (int j = threadIdx.x; j < K; j += blockDim.x) {
        const double f = freqs[j];
        const resp::NodeResp nr = resp::resp_node(orb, f);
        cd S[6];
        for (int k = 0; k < 6; k++) S[k] = mk(0.0, 0.0);
        for (int k = 0; k < M; k++) {
            const FusedRec& r = rb[ml.order[k]];
            double amp, phase, tf, ad;
            phm::phm_eval(r.w, f, amp, phase, tf, &ad);
            // phenomhm.py:487-507: phase += 2 pi f (t_ref - t0); tf += t_ref
            phase = phase + ((f * win.dt_shift) * 2) * PI;
            tf = tf + win.t_ref;
            cd G[6];
            double dphi;
            if (ml.pair[k]) {
                const FusedRec& r2 = rb[ml.order[k + 1]];
                const double Mf = f * r.w.M_sec;
                const bool same = Mf <= fmin(fmin(r.w.fi, r2.w.fi), fmin(r.w.amp_fi, r2.w.amp_fi)) &&
                                  r.w.ai * Mf < phm::PHI_JOIN_INS;
                if (__all(same)) {
                    const double amp2 = (ad * phm::hm_amp_factor(r2.w, Mf, r2.w.amp_Ai * Mf)) * r2.w.amp_scale;
                    const cd Yp = r.s.Yp * amp + r2.s.Yp * amp2, Yc = r.s.Yc * amp + r2.s.Yc * amp2;
                    resp::resp_links_y(r.s, Yp, Yc, orb, nr, f, tf, G, dphi, tab);
                    if (in_window(win, tf)) {
                        const cd e = cexpi_tab(phase + dphi, tab);
                        for (int q = 0; q < 6; q++) S[q] += G[q] * e;
                    }
                    k++;
                    continue;
                }
            }
            resp::resp_links(r.s, orb, nr, f, tf, G, dphi, tab);
            // The waveform phase multiplies the link terms after they are
            // formed, not inside their exponentials: the TDI combination takes
            // differences of nearly equal link terms at low frequency, so the
            // link terms' relative phases must not carry the rounding of this
            // large phase (1e4-1e7 rad). Modes outside the observation window
            // contribute nothing (WaveformBuild.cu:50).
            if (in_window(win, tf)) {
                const cd ap = amp * cexpi_tab(phase + dphi, tab);
                for (int k = 0; k < 6; k++) S[k] += G[k] * ap;
            }
        }
        cd T[3];
        resp::tdi_combine(S, nr, TDItag, rescaled, tdi2, T);
        // r = h / h0 with h = conj(T) (likelihood.py:603), as h (1 / h0)
        for (int c = 0; c < 3; c++) rs[c * K + j] = conj(T[c]) * nw[c * K + j].q;
    }
}

// The records of one template copied into LDS ('srec', M of them) by the
// threads of a workgroup; a barrier must follow. The residual pass then reads
// them with LDS loads (wave-uniform addresses, one broadcast each) instead of
// scalar loads: measured 140 -> 129 us for 4096 templates x 128 nodes (the
// scalar-load waits per mode, which the SGPR budget serialises, gone). The
// copy itself costs about 3 us of the kernel at that size (measured by doing
// it twice).
__device__ inline void records_to_lds(FusedRec* srec, const FusedRec* __restrict__ rb, int M)
{
    const double* src = reinterpret_cast<const double*>(rb);
    double* dst = reinterpret_cast<double*>(srec);
    for (int i = threadIdx.x; i < (int)(M * sizeof(FusedRec) / sizeof(double)); i += blockDim.x) dst[i] = src[i];
}

// The e^{i theta} table of cexpi_tab copied from constant memory into LDS
// ('tab', 129 entries) by the threads of a workgroup; a barrier must follow.
__device__ inline void etab_to_lds(double2* tab)
{
    for (int j = threadIdx.x; j < 129; j += blockDim.x) tab[j] = make_double2(ETAB[j].x, ETAB[j].y);
}

// Heterodyned sums (Likelihood.cu:52-111) in the per-node form of NodeW over
// the nodes j = t, t + blockDim.x, ... of thread t. Returns the workgroup
// totals (d_h re, im, h_h re, im) in every thread: each wavefront reduces its
// lanes, then the (at most 8) wavefront totals are added in wavefront order.
__device__ inline void node_sums(const cd* rs, const NodeW* __restrict__ nw, int K, double out[4])
{
    double l1r = 0.0, l1i = 0.0, l2r = 0.0, l2i = 0.0;
    for (int j = threadIdx.x; j < K; j += blockDim.x) {
        for (int c = 0; c < 3; c++) {
            const NodeW& w = nw[c * K + j];
            const cd r = rs[c * K + j];
            // conj(r) W
            l1r += fma(r.re, w.w.re, r.im * w.w.im);
            l1i += fma(r.re, w.w.im, -(r.im * w.w.re));
            const double m2 = fma(r.re, r.re, r.im * r.im);
            double x = 0.0;
            if (j > 0) {
                const cd p = rs[c * K + j - 1];
                x = fma(r.re, p.re, r.im * p.im);
            }
            l2r += fma(m2, w.v.re, x * w.u.re);
            l2i += fma(m2, w.v.im, x * w.u.im);
        }
    }
    out[0] = wsum(l1r);
    out[1] = wsum(l1i);
    out[2] = wsum(l2r);
    out[3] = wsum(l2i);
    const int nwave = blockDim.x >> 6;
    if (nwave > 1) {
        __shared__ double part[4][MAX_RESIDUAL_WAVES];
        if ((threadIdx.x & 63) == 0)
            for (int k = 0; k < 4; k++) part[k][threadIdx.x >> 6] = out[k];
        __syncthreads();
        for (int k = 0; k < 4; k++) {
            double s = part[k][0];
            for (int i = 1; i < nwave; i++) s += part[k][i];
            out[k] = s;
        }
    }
}

// Wavefronts per template of the residual pass (workgroup of 64 x that many
// threads): enough for about 2048 wavefronts in the launch (two per SIMD of
// MI355X's 1024), at most one per 64 nodes and at most MAX_RESIDUAL_WAVES.
// The forward and the gradient's weights pass use the same count, so their
// sums agree bit for bit.
inline int residual_waves(int B, int K)
{
    int nw = (2048 + B - 1) / B;
    nw = nw < (K + 63) / 64 ? nw : (K + 63) / 64;
    nw = nw < MAX_RESIDUAL_WAVES ? nw : MAX_RESIDUAL_WAVES;
    return nw > 1 ? nw : 1;
}

// Host helpers shared by the fused entry points (hdyn_fused.hip): three timing
// events when bbhx_hip_timing_enable(1) is on, and the (l, m) list check.
bool timer_take3(hipEvent_t ev[3]);
int check_modes(const int* ells_host, const int* mms_host, int numModes, ModeList& ml);

}  // namespace fused
}  // namespace bbhx
