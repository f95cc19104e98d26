// Derivatives of the waveform and response with respect to the 11 source
// parameters (ln M, q, chi1, chi2, d, phi_ref, cos inc, lam, sin beta, psi,
// t_ref), shared by the adjoint kernels of the heterodyned and the direct
// log-likelihood (no counterpart in the reference, which has no derivatives):
// records with tangents per (template, mode) and the per-point contraction of
// the waveform's and response's tangents with the likelihood's adjoint weights.
#pragma once

#include "common.h"
#include "dual.h"
#include "fused_device.h"

namespace bbhx {
namespace gradk {

constexpr int NI = 4;  // intrinsic tangent directions: ln M, q, chi1, chi2
using D4 = Dual<NI>;
constexpr int NGRAD = 11;
enum { G_LNM, G_Q, G_CHI1, G_CHI2, G_DIST, G_PHI, G_COSI, G_LAM, G_SINB, G_PSI, G_TREF };

// The frame (u, v, k) is orthonormal, so its derivative in any sky angle p is a
// rotation: dk/dp = a u + b v, du/dp = -a k + c v, dv/dp = -b k - c u with
// a = u.dk/dp, b = v.dk/dp, c = v.du/dp (psi: a = b = 0, c = 1). rot[q] holds
// (a, b, c) for p = lam (q = 0) and beta (q = 1); the derivative of any
// projection n.k, n.u, n.v then follows from the three projections.
struct GradResp {
    double u[3], v[3], k[3];
    double rot[2][3];     // (a, b, c) for lam, beta
    cd Yp, Yc, dYp, dYc;  // dY/d(cos inc)
    double m;             // azimuthal number
    double inv_cosb;      // d beta / d(sin beta)
};

struct GradRec {
    phm::PhmRecT<D4> w;
    GradResp s;
};

// Re(a b)
__device__ inline double re_mul(const cd& a, const cd& b) { return fma(a.re, b.re, -(a.im * b.im)); }

__device__ inline void grad_resp_setup(GradResp& r, int l, int m, double phi, double inc, double lam, double beta,
                                       double psi)
{
    const double cl = cos(lam), sl = sin(lam), cb = cos(beta), sb = sin(beta), cp = cos(psi), sp = sin(psi);
    r.k[0] = -cb * cl;
    r.k[1] = -cb * sl;
    r.k[2] = -sb;
    r.u[0] = cp * sl - cl * sb * sp;
    r.u[1] = -cl * cp - sb * sl * sp;
    r.u[2] = cb * sp;
    r.v[0] = -cl * cp * sb - sl * sp;
    r.v[1] = -cp * sb * sl + cl * sp;
    r.v[2] = cb * cp;
    // dk and du in lam (q = 0) and beta (q = 1), projected on the frame
    double dk[2][3], du[2][3];
    dk[0][0] = cb * sl;
    dk[0][1] = -cb * cl;
    dk[0][2] = 0.0;
    du[0][0] = cp * cl + sl * sb * sp;
    du[0][1] = sl * cp - sb * cl * sp;
    du[0][2] = 0.0;
    dk[1][0] = sb * cl;
    dk[1][1] = sb * sl;
    dk[1][2] = -cb;
    du[1][0] = -cl * cb * sp;
    du[1][1] = -cb * sl * sp;
    du[1][2] = -sb * sp;
    for (int q = 0; q < 2; q++) {
        r.rot[q][0] = resp::dot3(r.u, dk[q]);
        r.rot[q][1] = resp::dot3(r.v, dk[q]);
        r.rot[q][2] = resp::dot3(r.v, du[q]);
    }
    r.inv_cosb = 1.0 / cb;
    r.m = m;
    // Y+ and Yx as resp_setup forms them, and their derivatives in inc
    const Dual<1> th = Dual<1>::seed(inc, 0);
    const Dual<1> fp = resp::ylm_fac(l, m, th), fm = resp::ylm_fac(l, -m, th);
    const double sgn = (l & 1) ? -1.0 : 1.0;
    const cd e = cexpi(m * phi);
    const double a = 0.5 * (fp.v + sgn * fm.v), c = 0.5 * (fp.v - sgn * fm.v);
    const double da = 0.5 * (fp.d[0] + sgn * fm.d[0]), dc = 0.5 * (fp.d[0] - sgn * fm.d[0]);
    r.Yp = a * e;
    r.Yc = mul_i(c * e);
    // d/d(cos inc) = -(1 / sin inc) d/d inc
    const double si = sin(inc);
    const double dci = si != 0.0 ? -1.0 / si : 0.0;
    r.dYp = (da * dci) * e;
    r.dYc = mul_i((dc * dci) * e);
}


// The record with tangents of one (template, mode) (PhenomHMWaveform.cu set-up
// code instantiated on Dual<4>, seeded in ln M, q, chi1, chi2) and the
// response basis with its derivatives.
__device__ inline void grad_setup_one(GradRec& g, const double* __restrict__ params, int B, int b, int l, int m,
                                      const double* __restrict__ qnm, int table_len, double dspin)
{
    // masses from (ln M, q): m1 = M q / (1 + q), m2 = M / (1 + q)
    const double m1 = params[fused::P_M1 * B + b], m2 = params[fused::P_M2 * B + b];
    const double Mt = m1 + m2, q = m1 / m2;
    const double dq = Mt / ((1.0 + q) * (1.0 + q));
    D4 m1d(m1), m2d(m2);
    m1d.d[0] = m1;
    m1d.d[1] = dq;
    m2d.d[0] = m2;
    m2d.d[1] = -dq;
    const D4 c1d = D4::seed(params[fused::P_CHI1 * B + b], 2), c2d = D4::seed(params[fused::P_CHI2 * B + b], 3);
    D4 finmass, finspin, rdD, dmD, rd, dm;
    phm::final_state(m1d, m2d, c1d, c2d, finmass, finspin);
    fused::phenomd_ringdown(finmass, finspin, qnm, table_len, dspin, rdD, dmD);
    phm::hm_ringdown(l, m, finmass, finspin, rd, dm);
    phm::phm_setup(g.w, m1d * MSUN_SI, m2d * MSUN_SI, c1d, c2d, D4(params[fused::P_DIST * B + b]), l, m, rd, dm, rdD,
                   dmD, false);
    GradResp s;
    grad_resp_setup(s, l, m, params[fused::P_PHI * B + b], params[fused::P_INC * B + b],
                    params[fused::P_LAM * B + b], params[fused::P_BETA * B + b], params[fused::P_PSI * B + b]);
    g.s = s;
}

// ------------------------------------------------------------ PhenomHM tangents
//
// Amplitude, phase and t(f) of one (binary, mode) with their derivatives in the
// four intrinsic directions: the evaluation of phm::phm_eval<D4> (same
// branches, same expressions), arranged so that the functions of the mapped
// frequency x alone (its powers, logarithm, reciprocal) are evaluated once with
// their derivative in x (Dual<1>), and every sum of record coefficients times
// such functions is accumulated as value, coefficient tangents and x
// derivative, x's own tangents applied once at the end (chain rule):
//   d/dk sum_i c_i b_i(x) = sum_i c_i,k b_i(x) + x_k sum_i c_i b_i'(x).
// Only the pieces that are not of that form (the ringdown Lorentzian and
// arctangent, the mode's PN amplitude factor, the final combination) carry four
// tangents through every operation.

using X1 = Dual<1>;

struct Lin {
    double v, c[NI], x;
};

__device__ inline Lin lin0(double v = 0.0)
{
    Lin L;
    L.v = v;
    for (int k = 0; k < NI; k++) L.c[k] = 0.0;
    L.x = 0.0;
    return L;
}

// L += c b(x)
__device__ inline void lin(Lin& L, const D4& c, const X1& b)
{
    L.v = fma(c.v, b.v, L.v);
    for (int k = 0; k < NI; k++) L.c[k] = fma(c.d[k], b.v, L.c[k]);
    L.x = fma(c.v, b.d[0], L.x);
}

// L += c
__device__ inline void lin(Lin& L, const D4& c)
{
    L.v += c.v;
    for (int k = 0; k < NI; k++) L.c[k] += c.d[k];
}

// the accumulated sum as a D4, at the point x (value and tangents of x)
__device__ inline D4 lin_at(const Lin& L, const D4& x)
{
    D4 r;
    r.v = L.v;
    for (int k = 0; k < NI; k++) r.d[k] = fma(L.x, x.d[k], L.c[k]);
    return r;
}

// phm::amp_imr at x = Mf_amp (xa: its value and tangents; x: the same value as
// the Dual<1> variable; p: the powers of x)
__device__ inline D4 amp_imr_tan(const phm::PhmRecT<D4>& r, const D4& xa, const X1& x, const phm::P6T<X1>& p)
{
    Lin Lp = lin0();
    lin(Lp, r.a_amp0, rcp_pos(x * p.sixth));
    const D4 pre = lin_at(Lp, xa);
    if (!(x.v >= phm::AMP_JOIN_INS)) {
        Lin L = lin0(1.0);
        lin(L, r.a_pf23, p.two_thirds);
        lin(L, r.a_pf1, x);
        lin(L, r.a_pf43, p.four_thirds);
        lin(L, r.a_pf53, p.five_thirds);
        lin(L, r.a_pf2, p.two);
        lin(L, r.a_rho1, p.seven_thirds);
        lin(L, r.a_rho2, p.eight_thirds);
        lin(L, r.a_rho3, x * p.two);
        return pre * lin_at(L, xa);
    }
    if (x.v >= r.a_fmax.v) return pre * phm::amp_mrd(r, xa);
    const X1 x3 = x * p.two, x4 = x * x3;
    Lin L = lin0();
    lin(L, r.a_d0);
    lin(L, r.a_d1, x);
    lin(L, r.a_d2, p.two);
    lin(L, r.a_d3, x3);
    lin(L, r.a_d4, x4);
    return pre * lin_at(L, xa);
}

// phm::phase_imr at x = Mf_ph: the PhenomD phase and its derivative in Mf
__device__ inline void phase_imr_tan(const phm::PhmRecT<D4>& r, const D4& xp, const X1& x,
                                     const phm::P6T<X1>* pre, bool same, D4& ph, D4& dph)
{
    const X1 iMf = rcp_pos(x);
    const bool ins = !(x.v >= phm::PHI_JOIN_INS);
    if (!ins && x.v >= r.fMRDJoin.v) {
        const X1 f075 = sqrt(x * sqrt(x));
        const D4 q = (xp - r.alpha5 * r.p_fRD) * rcp_pos(r.p_Rholm * r.p_fDM * r.p_Taulm);
        Lin L = lin0();
        lin(L, r.alpha2, -iMf);
        lin(L, r.alpha3, (4.0 / 3.0) * f075);
        lin(L, r.alpha1, x);
        const D4 phi_mrd = lin_at(L, xp) + r.alpha4 * r.p_Rholm * atan_fast(q);
        Lin C = lin0();
        lin(C, r.C1MRD);
        lin(C, r.C2MRD, x);
        ph = r.p_etaInv * phi_mrd + lin_at(C, xp);
        Lin Ld = lin0();
        lin(Ld, r.alpha1);
        lin(Ld, r.alpha2, iMf * iMf);
        lin(Ld, r.alpha3, f075 * iMf);
        const D4 dm = lin_at(Ld, xp) + r.alpha4 * rcp_pos(r.p_fDM * r.p_Taulm * (1.0 + q * q));
        dph = dm * r.p_etaInv + r.C2MRD;
        return;
    }
    const X1 logMf = log_pos(x);
    if (!ins) {
        Lin L = lin0();
        lin(L, r.beta1, x);
        lin(L, r.beta3, (-1.0 / 3.0) * (iMf * iMf * iMf));
        lin(L, r.beta2, logMf);
        Lin C = lin0();
        lin(C, r.C1Int);
        lin(C, r.C2Int, x);
        ph = r.p_etaInv * lin_at(L, xp) + lin_at(C, xp);
        const X1 im2 = iMf * iMf;
        Lin Ld = lin0();
        lin(Ld, r.beta1);
        lin(Ld, r.beta3, im2 * im2);
        lin(Ld, r.beta2, iMf);
        dph = lin_at(Ld, xp) * r.p_etaInv + r.C2Int;
        return;
    }
    phm::P6T<X1> p;
    if (pre && same) p = *pre;
    else p = phm::pow6(x);
    const double pi13 = 1.4645918875615231, LOG_PI = 1.1447298858494002;
    const X1 v = p.third * pi13;
    const X1 logv = (logMf + LOG_PI) * (1.0 / 3.0);
    const X1 im13 = p.two_thirds * iMf;
    const X1 im23 = im13 * im13;
    Lin L = lin0();
    lin(L, r.q_init);
    lin(L, r.q_23, p.two_thirds);
    lin(L, r.q_13, p.third);
    lin(L, r.q_13log, logv * p.third);
    lin(L, r.q_logv, logv);
    lin(L, r.q_m13, im13);
    lin(L, r.q_m23, im23);
    lin(L, r.q_m1, iMf);
    lin(L, r.q_m53, im23 * iMf);
    Lin L2 = lin0();
    lin(L2, r.q_1, x);
    lin(L2, r.q_43, p.four_thirds);
    lin(L2, r.q_53, p.five_thirds);
    lin(L2, r.q_2, p.two);
    ph = lin_at(L, xp) + lin_at(L2, xp) * r.p_etaInv;
    // d phi_ins / dMf with v = (pi Mf)^(1/3) (phm::dphi_ins)
    const X1 v2 = v * v, v3 = v * v2, v4 = v * v3, v5 = v * v4, v6 = v * v5, v7 = v * v6;
    const X1 iv = rcp_pos(v), iv2 = iv * iv, iv4 = iv2 * iv2;
    const X1 w = (iv4 * iv4) * (PI / 3.0);
    Lin Ld = lin0();
    lin(Ld, r.pn_v7, 2.0 * (v7 * w));
    lin(Ld, r.pn_v6, v6 * w);
    lin(Ld, r.pn_vl6, (1.0 + logv) * (v6 * w));
    lin(Ld, r.pn_vl5, v5 * w);
    lin(Ld, r.pn_v4, -1.0 * (v4 * w));
    lin(Ld, r.pn_v3, -2.0 * (v3 * w));
    lin(Ld, r.pn_v2, -3.0 * (v2 * w));
    lin(Ld, r.pn_v0, -5.0 * w);
    const double ipi13 = 0.68278406325529568, ipi23 = 0.46619407703541166, ipi = 0.31830988618379067;
    Lin Ls = lin0();
    lin(Ls, r.sigma1);
    lin(Ls, r.sigma2, v * ipi13);
    lin(Ls, r.sigma3, v2 * ipi23);
    lin(Ls, r.sigma4, v3 * ipi);
    dph = lin_at(Ld, xp) + lin_at(Ls, xp) * r.p_etaInv;
}

// phm::pn_habs2: |H(v)|^2 with its polynomial parts as sums in v
__device__ inline D4 pn_habs2_tan(const phm::PhmRecT<D4>& r, const D4& v)
{
    const X1 x = X1::seed(v.v, 0);
    const X1 x2 = x * x, x3 = x * x2, x4 = x2 * x2;
    Lin Lr = lin0();
    lin(Lr, r.pn_c1, x);
    lin(Lr, r.pn_c2, x2);
    lin(Lr, r.pn_c3, x3);
    lin(Lr, r.pn_c4, x4);
    Lin Li = lin0();
    lin(Li, r.pn_c4i, x4);
    const D4 re = lin_at(Lr, v), im = lin_at(Li, v);
    return re * re + im * im;
}

// phm::hm_amp_factor
__device__ inline D4 hm_amp_factor_tan(const phm::PhmRecT<D4>& r, const D4& Mf_wf, const D4& Mf_amp)
{
    if (r.ell == 2 && r.mm == 2) return D4(1.0);
    const D4 v1 = cbrt_pos(r.pn_vscale * Mf_wf);
    const D4 h1 = pn_habs2_tan(r, v1);
    if (h1.v == 0.0) return D4(0.0);
    if (!(Mf_wf.v > r.amp_fi.v)) return sqrt(h1);
    const D4 v2 = v1 * r.pn_v2ratio;
    const D4 v3 = cbrt_pos(r.pn_vscale * Mf_amp);
    return sqrt(h1 * pn_habs2_tan(r, v3) * rcp_pos(pn_habs2_tan(r, v2)));
}

// phm::phm_eval<D4>: amplitude, phase and t(f) with their intrinsic tangents
__device__ inline void phm_eval_tan(const phm::PhmRecT<D4>& r, double f, D4& amp, D4& phase, D4& tf)
{
    const D4 Mf_wf = f * r.M_sec;
    D4 a, b;
    if (Mf_wf.v > r.amp_fi.v) {
        if (Mf_wf.v > r.amp_fr.v) {
            a = D4(1.0);
            b = r.amp_Br;
        } else {
            a = r.amp_Am;
            b = r.amp_Bm;
        }
    } else {
        a = r.amp_Ai;
        b = D4(0.0);
    }
    const D4 Mf_amp = a * Mf_wf + b;
    const X1 xa = X1::seed(Mf_amp.v, 0);
    const phm::P6T<X1> p = phm::pow6(xa);
    D4 A = amp_imr_tan(r, Mf_amp, xa, p);
    A *= hm_amp_factor_tan(r, Mf_wf, Mf_amp);
    amp = A * r.amp_scale;

    D4 sa, isa, sb, c1, c2;
    if (!(Mf_wf.v > r.fi.v)) {
        sa = r.ai;
        isa = r.iai;
        sb = r.bi;
        c1 = D4(0.0);
        c2 = D4(0.0);
    } else if (!(Mf_wf.v > r.fr.v)) {
        sa = r.am;
        isa = r.iam;
        sb = r.bm;
        c1 = r.PhDBconst;
        c2 = r.PhDBAterm;
    } else {
        sa = r.ar;
        isa = r.iar;
        sb = r.br;
        c1 = r.PhDCconst;
        c2 = r.phaseC;
    }
    const D4 Mf_ph = sa * Mf_wf + sb;
    D4 ph, dph;
    phase_imr_tan(r, Mf_ph, X1::seed(Mf_ph.v, 0), &p, Mf_ph.v == Mf_amp.v, ph, dph);
    const D4 phi = r.cshift + ((ph * isa - c1) + c2);
    phase = -r.t0 * (Mf_wf - r.Mf_ref) + (phi - (r.mm * r.phi0));
    const double INV_TWO_PI = 0.15915494309189535;
    tf = (r.M_sec * (dph * INV_TWO_PI)) - ((r.t0 * INV_TWO_PI) * r.M_sec);
}

// The same record in three independent parts (the parts of phm_setup, and the
// response basis), so that a set-up kernel can run each long dependency chain
// in a wavefront of its own: PART 0 the mode-dependent PhenomHM coefficients,
// PART 1 the (2,2) reference, PART 2 the response basis. They write disjoint
// fields.
template <int PART>
__device__ inline void grad_setup_part(GradRec& g, const double* __restrict__ params, int B, int b, int l, int m,
                                       const double* __restrict__ qnm, int table_len, double dspin)
{
    if constexpr (PART == 2) {
        GradResp s;
        grad_resp_setup(s, l, m, params[fused::P_PHI * B + b], params[fused::P_INC * B + b],
                        params[fused::P_LAM * B + b], params[fused::P_BETA * B + b], params[fused::P_PSI * B + b]);
        g.s = s;
    } else {
        const double m1 = params[fused::P_M1 * B + b], m2 = params[fused::P_M2 * B + b];
        const double Mt = m1 + m2, q = m1 / m2;
        const double dq = Mt / ((1.0 + q) * (1.0 + q));
        D4 m1d(m1), m2d(m2);
        m1d.d[0] = m1;
        m1d.d[1] = dq;
        m2d.d[0] = m2;
        m2d.d[1] = -dq;
        const D4 c1d = D4::seed(params[fused::P_CHI1 * B + b], 2), c2d = D4::seed(params[fused::P_CHI2 * B + b], 3);
        D4 finmass, finspin, rdD, dmD, rd, dm;
        phm::final_state(m1d, m2d, c1d, c2d, finmass, finspin);
        fused::phenomd_ringdown(finmass, finspin, qnm, table_len, dspin, rdD, dmD);
        if constexpr (PART == 0) phm::hm_ringdown(l, m, finmass, finspin, rd, dm);
        phm::phm_setup<D4, PART>(g.w, m1d * MSUN_SI, m2d * MSUN_SI, c1d, c2d, D4(params[fused::P_DIST * B + b]), l,
                                 m, rd, dm, rdD, dmD, false);
    }
}

__device__ inline double dsinc(double a, double cos_a, double sinc_a) { return a == 0.0 ? 0.0 : (cos_a - sinc_a) / a; }

// Gradient contributions of the response of one (frequency, mode) to the
// extrinsic entries of acc[11] (d, phi_ref, cos inc, lam, sin beta, psi,
// t_ref), for a log-likelihood whose first-order change is
//   d logL = a_A dA + a_phi d phi + Re sum_c W_c dT_c,
// where A is the amplitude, phi the phase (PhenomHM phase + 2 pi f (t_ref - t0)
// + the orbital Doppler phase) and T_c the TDI transfer functions at f and
// t = t(f) + t_ref. weights(T, E, A, W, a_A, a_phi) supplies W, a_A, a_phi once
// T, E = e^{i phi} and A are known. Av, Phv, Tfv: the PhenomHM amplitude, phase
// and t(f) at f. On return (true) the intrinsic parameters' contribution is
//   a_t dt(f) + a_A dA + a_phi dPhi   (a_t = dlogL/dt at fixed phase),
// so the three scalars are all the intrinsic tangents need. use_window: return
// false, contributing nothing, if t lies outside the observation window (the
// heterodyned form's rule for its nodes).
template <class Weights, class E>
__device__ inline bool grad_resp_part(const GradResp& R, const resp::OrbitTab& o, double f, double Av, double Phv,
                                      double Tfv, const fused::Window& win, bool use_window, Weights weights,
                                      int TDItag, int rescaled, int tdi2, double dist, double acc[NGRAD],
                                      const E* __restrict__ tab, double& a_t, double& a_A, double& a_phi)
{
    auto sincos_t = [&](double th, double* s_, double* c_) {
        const cd e = cexpi_tab(th, tab);
        *s_ = e.im;
        *c_ = e.re;
    };
    const double t = Tfv + win.t_ref;
    if (use_window && !fused::in_window(win, t)) return false;

    // orbit projections on the frame and their time derivatives, from the
    // packed row of the table (values P and differences D to the next row:
    // P(t) = P + a D, dP/dt = D / dt); derivatives in the sky angles follow
    // from the projections (the frame's rotation)
    int w = int(floor(t * o.inv_dt));
    w = w < 0 ? 0 : (w > o.N - 2 ? o.N - 2 : w);
    const double a = (t - o.dt * w) * o.inv_dt;
    const double* row = o.p + (long)w * resp::ORB_ROW;
    // the interpolated vector q (0..2 spacecraft, 3..5 links 12, 23, 31) and its difference
    auto vec = [&](int q, double v[3], double d[3]) {
        for (int c = 0; c < 3; c++) {
            d[c] = row[resp::ORB_VALS + 3 * q + c];
            v[c] = fma(d[c], a, row[3 * q + c]);
        }
    };
    double kp[3], dkp_t[3], up[3], vp[3];
    for (int i = 0; i < 3; i++) {
        double v[3], d[3];
        vec(i, v, d);
        kp[i] = resp::dot3(R.k, v);
        dkp_t[i] = resp::dot3(R.k, d) * o.inv_dt;
        up[i] = resp::dot3(R.u, v);
        vp[i] = resp::dot3(R.v, v);
    }
    double kn[3], un[3], vn[3], dkn_t[3], dun_t[3], dvn_t[3];
    for (int j = 0; j < 3; j++) {
        double v[3], d[3];
        vec(3 + j, v, d);
        kn[j] = resp::dot3(R.k, v);
        dkn_t[j] = resp::dot3(R.k, d) * o.inv_dt;
        un[j] = resp::dot3(R.u, v);
        dun_t[j] = resp::dot3(R.u, d) * o.inv_dt;
        vn[j] = resp::dot3(R.v, v);
        dvn_t[j] = resp::dot3(R.v, d) * o.inv_dt;
    }

    // orbital Doppler phase and its derivatives (Response.cu:278-296)
    double sa, ca;
    sincos_t(OMEGA0 * t, &sa, &ca);
    const double c2f = 2. * PI / C_SI * f;
    const double pRd = c2f * (R.k[0] * (AU_SI * ca) + R.k[1] * (AU_SI * sa));
    const double pRd_t = c2f * AU_SI * OMEGA0 * (-R.k[0] * sa + R.k[1] * ca);
    // the Doppler phase's derivative in a sky angle is c2f (a u.R + b v.R)
    const double pRd_u = c2f * AU_SI * (R.u[0] * ca + R.u[1] * sa);
    const double pRd_v = c2f * AU_SI * (R.v[0] * ca + R.v[1] * sa);
    const double phase = Phv + ((f * win.dt_shift) * 2) * PI + pRd;

    // single-link terms as resp_eval forms them
    const double x = PI * f * o.L / C_SI;
    double sx, cx;
    sincos_t(x, &sx, &cx);
    cd nHn[3];
    double sm[3], sp[3], dsm[3], dsp[3];
    for (int j = 0; j < 3; j++) {
        nHn[j] = R.Yp * (un[j] * un[j] - vn[j] * vn[j]) + R.Yc * (2.0 * un[j] * vn[j]);
        double s, c;
        sincos_t(x * kn[j], &s, &c);
        const double am = x * (1. - kn[j]), ap = x * (1. + kn[j]);
        const double sin_m = sx * c - cx * s, sin_p = sx * c + cx * s;
        const double cos_m = cx * c + sx * s, cos_p = cx * c - sx * s;
        const double prod = am * ap;
        if (prod != 0.0) {  // 1/am and 1/ap from one reciprocal of their product
            const double ip = rcp_pos(prod);
            const double iam = ap * ip, iap = am * ip;
            sm[j] = sin_m * iam;
            sp[j] = sin_p * iap;
            dsm[j] = (cos_m - sm[j]) * iam;
            dsp[j] = (cos_p - sp[j]) * iap;
        } else {  // f = 0 or a link along k: sinc(0) = 1, sinc'(0) = 0
            sm[j] = am == 0.0 ? 1.0 : sin_m / am;
            sp[j] = ap == 0.0 ? 1.0 : sin_p / ap;
            dsm[j] = dsinc(am, cos_m, sm[j]);
            dsp[j] = dsinc(ap, cos_p, sp[j]);
        }
    }
    // link L = 0: pair (23, 32) with e23; L = 1: (31, 13) with e31; L = 2: (12, 21) with e12
    const int ka[3] = {1, 2, 0}, kb[3] = {2, 0, 1};
    cd eL[3], gL[3];
    for (int L = 0; L < 3; L++) {
        eL[L] = cexpi_tab(x * (1. + (kp[ka[L]] + kp[kb[L]]) * o.inv_L), tab);
        gL[L] = mul_i(x * nHn[L]);
    }
    // Gm = g sm e (G23, G31, G12), Gp = g sp e (G32, G13, G21)
    cd Gm[3], Gp[3];
    for (int L = 0; L < 3; L++) {
        Gm[L] = (gL[L] * sm[L]) * eL[L];
        Gp[L] = (gL[L] * sp[L]) * eL[L];
    }
    const cd &G23 = Gm[0], &G32 = Gp[0], &G31 = Gm[1], &G13 = Gp[1], &G12 = Gm[2], &G21 = Gp[2];

    // TDI combination (Response.cu:197-261) and its adjoint weights beta
    const double s2x = 2.0 * sx * cx, c2x = cx * cx - sx * sx;
    const cd z = mk(c2x, s2x);
    const cd one = mk(1.0, 0.0);
    cd T[3];
    cd bm[3], bp[3];  // beta of Gm[L], Gp[L]
    const cd E1 = cexpi_tab(phase, tab);
    cd W[3];
    double aA = 0.0, aPh = 0.0;
    if (TDItag == 1) {
        const cd fac = rescaled ? one : mul_i((2. * s2x) * z);
        T[0] = fac * (G21 + z * G12 - G31 - z * G13);
        T[1] = fac * (G32 + z * G23 - G12 - z * G21);
        T[2] = fac * (G13 + z * G31 - G23 - z * G32);
        weights(T, E1, Av, W, aA, aPh);
        const cd wX = W[0] * fac, wY = W[1] * fac, wZ = W[2] * fac;
        bp[2] = wX - z * wY;    // G21
        bm[2] = z * wX - wY;    // G12
        bm[1] = z * wZ - wX;    // G31
        bp[1] = wZ - z * wX;    // G13
        bp[0] = wY - z * wZ;    // G32
        bm[0] = z * wY - wZ;    // G23
    } else {
        cd fAE = one, fT = one;
        if (!rescaled) {
            fAE = mul_i((SQRT2 * s2x) * z);
            const double c3x = c2x * cx - s2x * sx, s3x = s2x * cx + c2x * sx;
            fT = (2. * SQRT2 * s2x * sx) * mk(c3x, s3x);
        }
        cd f2 = one;
        if (tdi2) {
            const double s4x = 2.0 * s2x * c2x, c4x = c2x * c2x - s2x * s2x;
            f2 = mul_i((-2. * s4x) * mk(c4x, s4x));
        }
        const cd fa = 2.0 * (f2 * fAE), ft = 2.0 * (f2 * fT);
        const cd Ac = 0.5 * ((one + z) * (G31 + G13) - G23 - z * G32 - G21 - z * G12);
        const cd Ec = (0.5 * INVSQRT3) * ((one - z) * (G13 - G31) + (mk(2.0, 0.0) + z) * (G12 - G32) +
                                         (one + 2. * z) * (G21 - G23));
        const cd Tc = INVSQRT6 * (G21 - G12 + G32 - G23 + G13 - G31);
        T[0] = fa * Ac;
        T[1] = fa * Ec;
        T[2] = ft * Tc;
        weights(T, E1, Av, W, aA, aPh);
        const cd wA = (W[0] * fa) * 0.5, wE = (W[1] * fa) * (0.5 * INVSQRT3), wT = (W[2] * ft) * INVSQRT6;
        bm[1] = wA * (one + z) - wE * (one - z) - wT;        // G31
        bp[1] = wA * (one + z) + wE * (one - z) + wT;        // G13
        bm[0] = mk(0.0, 0.0) - wA - wE * (one + 2. * z) - wT;  // G23
        bp[0] = mk(0.0, 0.0) - wA * z - wE * (mk(2.0, 0.0) + z) + wT;  // G32
        bp[2] = mk(0.0, 0.0) - wA + wE * (one + 2. * z) + wT;  // G21
        bm[2] = mk(0.0, 0.0) - wA * z + wE * (mk(2.0, 0.0) + z) - wT;  // G12
    }

    // sum_ij beta_ij dG_ij in t (Qt) and, through the projections, in the
    // sky angles and the Y factors. Per link L with G = g s e (g = i x nHn,
    // nHn = Yp U + Yc V, U = un^2 - vn^2, V = 2 un vn):
    //   dG = ixeB dnHn + egC dkn + egB (dkp_ka + dkp_kb)
    // (ixeB = i x e B, egC = e g C x, egB = e g B i x / L with B, C the sinc
    // terms' adjoint sums). With the frame's rotation (a, b, c) a sky angle
    // gives dU = 2 kn (b vn - a un) + 2 c V, dV = -2 kn (a vn + b un) - 2 c U,
    // dkn = a un + b vn, dkp = a up + b vp, so every sky derivative is
    //   a Sa + b Sb + c Spsi
    // with Spsi the psi derivative (a = b = 0, c = 1) and Sa, Sb sums over the
    // links; cos inc enters only through (Yp, Yc).
    cd Qt = mk(0.0, 0.0);
    cd SU = mk(0.0, 0.0), SV = mk(0.0, 0.0), Sku = mk(0.0, 0.0), Skv = mk(0.0, 0.0);
    cd Tu = mk(0.0, 0.0), Tv = mk(0.0, 0.0);
    const cd ix = mk(0.0, x);
    for (int L = 0; L < 3; L++) {
        const cd BL = bm[L] * sm[L] + bp[L] * sp[L];
        const cd CL = bp[L] * dsp[L] - bm[L] * dsm[L];
        const cd eB = eL[L] * BL;
        const cd eg = eL[L] * gL[L];
        const cd egC = eg * CL * x;
        const cd egB = eg * BL * ix * o.inv_L;
        const cd ixeB = ix * eB;
        const double U = un[L] * un[L] - vn[L] * vn[L], V = 2.0 * un[L] * vn[L];
        const double dU = 2.0 * (un[L] * dun_t[L] - vn[L] * dvn_t[L]);
        const double dV = 2.0 * (dun_t[L] * vn[L] + un[L] * dvn_t[L]);
        Qt += ixeB * (R.Yp * dU + R.Yc * dV) + egC * dkn_t[L] + egB * (dkp_t[ka[L]] + dkp_t[kb[L]]);
        SU += ixeB * U;
        SV += ixeB * V;
        Sku += ixeB * (kn[L] * un[L]);
        Skv += ixeB * (kn[L] * vn[L]);
        Tu += egC * un[L] + egB * (up[ka[L]] + up[kb[L]]);
        Tv += egC * vn[L] + egB * (vp[ka[L]] + vp[kb[L]]);
    }
    // Re of: Spsi = 2 (Yp SV - Yc SU), Sa = Tu - 2 (Yp Sku + Yc Skv),
    // Sb = Tv + 2 (Yp Skv - Yc Sku), Scos = dYp SU + dYc SV
    const double Spsi = 2.0 * (re_mul(R.Yp, SV) - re_mul(R.Yc, SU));
    const double Sa = Tu.re - 2.0 * (re_mul(R.Yp, Sku) + re_mul(R.Yc, Skv)) + aPh * pRd_u;
    const double Sb = Tv.re + 2.0 * (re_mul(R.Yp, Skv) - re_mul(R.Yc, Sku)) + aPh * pRd_v;
    const double Scos = re_mul(R.dYp, SU) + re_mul(R.dYc, SV);

    cd WT = mk(0.0, 0.0);
    for (int c = 0; c < 3; c++) WT += W[c] * T[c];
    a_t = Qt.re + aPh * pRd_t;
    a_A = aA;
    a_phi = aPh;
    acc[G_DIST] += -aA * Av / dist;
    acc[G_PHI] += -R.m * WT.im;  // T ~ e^{i m phi_ref}
    acc[G_COSI] += Scos;
    acc[G_LAM] += fma(R.rot[0][0], Sa, fma(R.rot[0][1], Sb, R.rot[0][2] * Spsi));
    acc[G_SINB] += fma(R.rot[1][0], Sa, fma(R.rot[1][1], Sb, R.rot[1][2] * Spsi)) * R.inv_cosb;
    acc[G_PSI] += Spsi;
    acc[G_TREF] += Qt.re + aPh * (2. * PI * f + pRd_t);
    return true;
}

// Both parts at one (frequency, mode): the waveform's values from the record
// with tangents.
template <class Weights, class E>
__device__ inline void grad_point_core(const GradRec& G, const resp::OrbitTab& o, double f, const fused::Window& win,
                                       bool use_window, Weights weights, int TDItag, int rescaled, int tdi2,
                                       double dist, double acc[NGRAD], const E* __restrict__ tab)
{
    D4 A, Ph, Tf;
    phm_eval_tan(G.w, f, A, Ph, Tf);
    double a_t, a_A, a_phi;
    if (!grad_resp_part(G.s, o, f, A.v, Ph.v, Tf.v, win, use_window, weights, TDItag, rescaled, tdi2, dist, acc, tab,
                        a_t, a_A, a_phi))
        return;
    for (int k = 0; k < NI; k++) acc[k] += a_t * Tf.d[k] + a_A * A.d[k] + a_phi * Ph.d[k];
}


}  // namespace gradk
}  // namespace bbhx
