// Gradient of the heterodyned log-likelihood with respect to the 11 source
// parameters (ln M, q, chi1, chi2, d, phi_ref, cos inc, lam, sin beta, psi, t_ref).
//
// There is no counterpart in the reference (it has no derivatives). The value
// differentiated is the one HeterodynedLikelihood.get_ll returns
// (likelihood.py:550-648): logL = Re d_h - 1/2 Re h_h - 1/2 d_d.
//
// k_grad_setup_fwd one lane per (template, mode): the double record of the
//                forward pass;
// k_grad_setup   (on a side stream, concurrently with the two kernels that
//                follow) one lane per (template, mode): a PhenomHM record whose
//                every coefficient carries its derivatives with respect to
//                (ln M, q, chi1, chi2) (Dual<4> instantiation of the set-up
//                code), plus the polarisation basis, wave vector and Y factors
//                with their derivatives in lam, beta (analytic) and cos inc.
// k_grad_weights one 64-lane wavefront per template:
//   1. forward: residuals r = h / h0 on the K nodes (LDS) and d_h, h_h,
//      exactly as the forward kernel;
//   2. adjoint of the heterodyned sums: for each node and channel the complex
//      weight g with  d logL = sum Re(g dh)  (from the segment formulas
//      of Likelihood.cu:91-103 and r = h / h0), written out (48 K bytes);
// k_grad_points  one 64-lane wavefront per (template, mode), lane per node:
//   3. per (node, mode) tangents: PhenomHM amplitude, phase and t(f) with
//      respect to the intrinsic parameters (Dual<4> evaluation of the same
//      phm_eval code), the response by an adjoint of the (linear) TDI
//      combination and forward tangents of the six single-link terms
//      G_ij in t, lam, beta, psi and cos inc; contracted with g at once, so
//      each lane keeps 11 accumulators; the wavefront reduces them;
// k_grad_reduce  the modes' partial gradients, summed in mode order.
// Splitting the forward pass from the per-point tangents keeps each kernel's
// register footprint (and so its occupancy) to its own needs.
// The spline and response are recomputed, nothing per node is stored.

#include "bbhx_hip.h"
#include "common.h"
#include "dual.h"
#include "fused_device.h"
#include "workspace.h"

using namespace bbhx;
using namespace bbhx::fused;

namespace bbhx {
bool response_orbit_tab(const bbhx_response* r, resp::OrbitTab& o);
}

namespace {

constexpr int NI = 4;  // intrinsic tangent directions: ln M, q, chi1, chi2
using D4 = Dual<NI>;
constexpr int NGRAD = 11;
enum { G_LNM, G_Q, G_CHI1, G_CHI2, G_DIST, G_PHI, G_COSI, G_LAM, G_SINB, G_PSI, G_TREF };

struct GradResp {
    double u[3], v[3], k[3];
    double dk[2][3], du[2][3], dv[2][3];  // d/d lam, d/d beta
    cd Yp, Yc, dYp, dYc;                  // dY/d(cos inc)
    double m;                             // azimuthal number
    double inv_cosb;                      // d beta / d(sin beta)
};

struct GradRec {
    phm::PhmRecT<D4> w;
    GradResp s;
};

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
    // d/d lam
    r.dk[0][0] = cb * sl;
    r.dk[0][1] = -cb * cl;
    r.dk[0][2] = 0.0;
    r.du[0][0] = cp * cl + sl * sb * sp;
    r.du[0][1] = sl * cp - sb * cl * sp;
    r.du[0][2] = 0.0;
    r.dv[0][0] = sl * cp * sb - cl * sp;
    r.dv[0][1] = -cp * sb * cl - sl * sp;
    r.dv[0][2] = 0.0;
    // d/d beta
    r.dk[1][0] = sb * cl;
    r.dk[1][1] = sb * sl;
    r.dk[1][2] = -cb;
    r.du[1][0] = -cl * cb * sp;
    r.du[1][1] = -cb * sl * sp;
    r.du[1][2] = -sb * sp;
    r.dv[1][0] = -cl * cp * cb;
    r.dv[1][1] = -cp * cb * sl;
    r.dv[1][2] = -sb * cp;
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

// the double records of the forward pass, one lane per (template, mode)
__global__ __launch_bounds__(64) void k_grad_setup_fwd(FusedRec* __restrict__ recs, const double* __restrict__ params,
                                                       int B, ModeList modes, int M, const double* __restrict__ qnm,
                                                       int table_len, double dspin)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * M) return;
    const int b = idx / M, mi = idx % M;
    setup_one(recs[idx], params, B, b, modes.ell[mi], modes.mm[mi], qnm, table_len, dspin);
}

// the records with tangents, one lane per (template, mode)
__global__ __launch_bounds__(64) void k_grad_setup(GradRec* __restrict__ grecs, const double* __restrict__ params,
                                                   int B, ModeList modes, int M, const double* __restrict__ qnm,
                                                   int table_len, double dspin)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * M) return;
    const int b = idx / M, mi = idx % M;
    const int l = modes.ell[mi], m = modes.mm[mi];
    // masses from (ln M, q): m1 = M q / (1 + q), m2 = M / (1 + q)
    const double m1 = params[P_M1 * B + b], m2 = params[P_M2 * B + b];
    const double Mt = m1 + m2, q = m1 / m2;
    const double dq = Mt / ((1.0 + q) * (1.0 + q));
    D4 m1d(m1), m2d(m2);
    m1d.d[0] = m1;
    m1d.d[1] = dq;
    m2d.d[0] = m2;
    m2d.d[1] = -dq;
    const D4 c1d = D4::seed(params[P_CHI1 * B + b], 2), c2d = D4::seed(params[P_CHI2 * B + b], 3);
    D4 finmass, finspin, rdD, dmD, rd, dm;
    phm::final_state(m1d, m2d, c1d, c2d, finmass, finspin);
    phenomd_ringdown(finmass, finspin, qnm, table_len, dspin, rdD, dmD);
    phm::hm_ringdown(l, m, finmass, finspin, rd, dm);
    GradRec& g = grecs[idx];
    phm::phm_setup(g.w, m1d * MSUN_SI, m2d * MSUN_SI, c1d, c2d, D4(params[P_DIST * B + b]), l, m, rd, dm, rdD, dmD,
                   false);
    GradResp s;
    grad_resp_setup(s, l, m, params[P_PHI * B + b], params[P_INC * B + b], params[P_LAM * B + b],
                    params[P_BETA * B + b], params[P_PSI * B + b]);
    g.s = s;
}

// Adjoint weights of the heterodyned sums: for node j and channel c,
// g[c][j] = gamma / conj(h0) with d logL = sum_j,c Re(conj(gamma) dr) and
// r = h / h0, so that d logL = sum Re(g dh) for h = sum_m conj(X_m).
__device__ inline void adjoint_weights(cd* gs, const cd* rs, const double* __restrict__ freqs,
                                       const cd* __restrict__ h0, const cd* __restrict__ dc, int K, int jend)
{
    for (int j = threadIdx.x; j < K; j += 64) {
        cd gam[3] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
        // segment s joins prev = s-1 and cur = s; node j is cur of s = j and prev of s = j+1
        for (int side = 0; side < 2; side++) {
            const int sgm = j + side;
            if (sgm < 1 || sgm >= jend) continue;
            const double fj = freqs[sgm], fprev = freqs[sgm - 1];
            if (fprev == 0.0) continue;
            const double inv = 1.0 / (fj - fprev);
            const double fm = (fj + fprev) / 2.0;
            const double w0 = side == 0 ? 1.0 - inv * (fj - fm) : inv * (fj - fm);  // d r0 / d r_j
            const double w1 = side == 0 ? inv : -inv;                                // d r1 / d r_j
            for (int c = 0; c < 3; c++) {
                const cd cur = rs[c * K + sgm], prev = rs[c * K + sgm - 1];
                const cd A0 = dc[(0 * 3 + c) * K + sgm], A1 = dc[(1 * 3 + c) * K + sgm];
                const double B0 = dc[(2 * 3 + c) * K + sgm].re, B1 = dc[(3 * 3 + c) * K + sgm].re;
                const cd r1 = mk((cur.re - prev.re) * inv, (cur.im - prev.im) * inv);
                const cd r0 = cur - r1 * (fj - fm);
                // d logL = Re(conj(g0) dr0 + conj(g1) dr1)
                const cd g0 = A0 - B0 * r0 - B1 * r1;
                const cd g1 = A1 - B1 * r0;
                gam[c] += w0 * g0 + w1 * g1;
            }
        }
        for (int c = 0; c < 3; c++) {
            const cd d = h0[c * K + j];
            const double iden = rcp_pos(d.re * d.re + d.im * d.im);
            gs[c * K + j] = mk((gam[c].re * d.re - gam[c].im * d.im) * iden, (gam[c].re * d.im + gam[c].im * d.re) * iden);
        }
    }
}

// derivative of sinc at a (sinc' = (cos a - sinc a) / a; 0 at a = 0)
__device__ inline double dsinc(double a, double cos_a, double sinc_a) { return a == 0.0 ? 0.0 : (cos_a - sinc_a) / a; }

// Gradient contributions of one (node, mode) to acc[11]. g: adjoint weights of
// the node's three channels.
__device__ inline void grad_point(const GradRec& G, const resp::OrbitTab& o, double f, const Window& win,
                                  const cd g[3], int TDItag, int rescaled, int tdi2, double dist, double acc[NGRAD])
{
    D4 A, Ph, Tf;
    phm::phm_eval(G.w, f, A, Ph, Tf);
    const double t = Tf.v + win.t_ref;
    if (!in_window(win, t)) return;
    const GradResp& R = G.s;

    // orbit projections (values), their time derivatives, and the spacecraft
    // projections' derivatives in lam, beta; the links' derivatives in lam,
    // beta are formed later, link by link, to keep fewer values live
    int w = int(floor(t * o.inv_dt));
    w = w < 0 ? 0 : (w > o.N - 2 ? o.N - 2 : w);
    const double a = (t - o.dt * w) * o.inv_dt;
    const double* x0 = o.x + (long)w * 9;
    const double* n0 = o.n + (long)w * 18;
    const int li[3] = {o.li12, o.li23, o.li31};
    double kp[3], dkp[3][3];
    for (int i = 0; i < 3; i++) {
        const double* lo = x0 + 3 * i;
        const double* hi = lo + 9;
        const double l0 = resp::dot3(R.k, lo), h0v = resp::dot3(R.k, hi);
        kp[i] = fma(h0v - l0, a, l0);
        dkp[0][i] = (h0v - l0) * o.inv_dt;
        for (int q = 0; q < 2; q++) {
            const double lq = resp::dot3(R.dk[q], lo), hq = resp::dot3(R.dk[q], hi);
            dkp[1 + q][i] = fma(hq - lq, a, lq);
        }
    }
    double kn[3], un[3], vn[3], dkn_t[3], dun_t[3], dvn_t[3];
    for (int j = 0; j < 3; j++) {
        const double* lo = n0 + 3 * li[j];
        const double* hi = lo + 18;
        double l0 = resp::dot3(R.k, lo), h0v = resp::dot3(R.k, hi);
        kn[j] = fma(h0v - l0, a, l0);
        dkn_t[j] = (h0v - l0) * o.inv_dt;
        l0 = resp::dot3(R.u, lo);
        h0v = resp::dot3(R.u, hi);
        un[j] = fma(h0v - l0, a, l0);
        dun_t[j] = (h0v - l0) * o.inv_dt;
        l0 = resp::dot3(R.v, lo);
        h0v = resp::dot3(R.v, hi);
        vn[j] = fma(h0v - l0, a, l0);
        dvn_t[j] = (h0v - l0) * o.inv_dt;
    }

    // orbital Doppler phase and its derivatives (Response.cu:278-296)
    double sa, ca;
    sincos_red(OMEGA0 * t, &sa, &ca);
    const double c2f = 2. * PI / C_SI * f;
    const double pRd = c2f * (R.k[0] * (AU_SI * ca) + R.k[1] * (AU_SI * sa));
    const double pRd_t = c2f * AU_SI * OMEGA0 * (-R.k[0] * sa + R.k[1] * ca);
    const double pRd_l = c2f * AU_SI * (R.dk[0][0] * ca + R.dk[0][1] * sa);
    const double pRd_b = c2f * AU_SI * (R.dk[1][0] * ca + R.dk[1][1] * sa);
    const double phase = Ph.v + ((f * win.dt_shift) * 2) * PI + pRd;

    // single-link terms as resp_eval forms them
    const double x = PI * f * o.L / C_SI;
    double sx, cx;
    sincos_red(x, &sx, &cx);
    cd nHn[3];
    double sm[3], sp[3], dsm[3], dsp[3];
    for (int j = 0; j < 3; j++) {
        nHn[j] = R.Yp * (un[j] * un[j] - vn[j] * vn[j]) + R.Yc * (2.0 * un[j] * vn[j]);
        double s, c;
        sincos_red(x * kn[j], &s, &c);
        const double am = x * (1. - kn[j]), ap = x * (1. + kn[j]);
        const double sin_m = sx * c - cx * s, sin_p = sx * c + cx * s;
        const double cos_m = cx * c + sx * s, cos_p = cx * c - sx * s;
        sm[j] = am == 0.0 ? 1.0 : sin_m / am;
        sp[j] = ap == 0.0 ? 1.0 : sin_p / ap;
        dsm[j] = dsinc(am, cos_m, sm[j]);
        dsp[j] = dsinc(ap, cos_p, sp[j]);
    }
    // link L = 0: pair (23, 32) with e23; L = 1: (31, 13) with e31; L = 2: (12, 21) with e12
    const int ka[3] = {1, 2, 0}, kb[3] = {2, 0, 1};
    cd eL[3], gL[3];
    for (int L = 0; L < 3; L++) {
        eL[L] = cexpi_red(x * (1. + (kp[ka[L]] + kp[kb[L]]) * o.inv_L));
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
    const cd E1 = cexpi_red(phase);
    const double Av = A.v;
    cd W[3];
    cd S1 = mk(0.0, 0.0);
    if (TDItag == 1) {
        const cd fac = rescaled ? one : mul_i((2. * s2x) * z);
        T[0] = fac * (G21 + z * G12 - G31 - z * G13);
        T[1] = fac * (G32 + z * G23 - G12 - z * G21);
        T[2] = fac * (G13 + z * G31 - G23 - z * G32);
        for (int c = 0; c < 3; c++) {
            S1 += g[c] * T[c] * E1;
            W[c] = (g[c] * E1) * Av;
        }
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
        for (int c = 0; c < 3; c++) {
            S1 += g[c] * T[c] * E1;
            W[c] = (g[c] * E1) * Av;
        }
        const cd wA = (W[0] * fa) * 0.5, wE = (W[1] * fa) * (0.5 * INVSQRT3), wT = (W[2] * ft) * INVSQRT6;
        bm[1] = wA * (one + z) - wE * (one - z) - wT;        // G31
        bp[1] = wA * (one + z) + wE * (one - z) + wT;        // G13
        bm[0] = mk(0.0, 0.0) - wA - wE * (one + 2. * z) - wT;  // G23
        bp[0] = mk(0.0, 0.0) - wA * z - wE * (mk(2.0, 0.0) + z) + wT;  // G32
        bp[2] = mk(0.0, 0.0) - wA + wE * (one + 2. * z) + wT;  // G21
        bm[2] = mk(0.0, 0.0) - wA * z + wE * (mk(2.0, 0.0) + z) - wT;  // G12
    }

    // Q_d = sum_ij beta_ij dG_ij for d = t, lam, beta (0..2), psi (3), cos inc (4)
    cd Q[5] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
    const cd ix = mk(0.0, x);
    for (int L = 0; L < 3; L++) {
        // this link's projections differentiated in t (0), lam (1), beta (2)
        double dknL[3], dunL[3], dvnL[3];
        dknL[0] = dkn_t[L];
        dunL[0] = dun_t[L];
        dvnL[0] = dvn_t[L];
        {
            const double* lo = n0 + 3 * li[L];
            const double* hi = lo + 18;
            for (int q = 0; q < 2; q++) {
                double lq = resp::dot3(R.dk[q], lo), hq = resp::dot3(R.dk[q], hi);
                dknL[1 + q] = fma(hq - lq, a, lq);
                lq = resp::dot3(R.du[q], lo);
                hq = resp::dot3(R.du[q], hi);
                dunL[1 + q] = fma(hq - lq, a, lq);
                lq = resp::dot3(R.dv[q], lo);
                hq = resp::dot3(R.dv[q], hi);
                dvnL[1 + q] = fma(hq - lq, a, lq);
            }
        }
        const cd BL = bm[L] * sm[L] + bp[L] * sp[L];
        const cd CL = bp[L] * dsp[L] - bm[L] * dsm[L];
        const cd eB = eL[L] * BL;
        const cd eg = eL[L] * gL[L];
        const cd egC = eg * CL * x;
        const cd egB = eg * BL * ix * o.inv_L;
        const cd ixeB = ix * eB;
        for (int d = 0; d < 3; d++) {
            const double dU = 2.0 * (un[L] * dunL[d] - vn[L] * dvnL[d]);
            const double dV = 2.0 * (dunL[d] * vn[L] + un[L] * dvnL[d]);
            const cd dnHn = R.Yp * dU + R.Yc * dV;
            Q[d] += ixeB * dnHn + egC * dknL[d] + egB * (dkp[d][ka[L]] + dkp[d][kb[L]]);
        }
        // psi: du = v, dv = -u
        {
            const double dU = 4.0 * un[L] * vn[L];
            const double dV = 2.0 * (vn[L] * vn[L] - un[L] * un[L]);
            Q[3] += ixeB * (R.Yp * dU + R.Yc * dV);
        }
        // cos inc: only the Y factors
        Q[4] += ixeB * (R.dYp * (un[L] * un[L] - vn[L] * vn[L]) + R.dYc * (2.0 * un[L] * vn[L]));
    }

    const double ReS = S1.re, ImS = S1.im;
    for (int k = 0; k < NI; k++)
        acc[k] += Q[0].re * Tf.d[k] + ReS * A.d[k] - ImS * Av * (Ph.d[k] + pRd_t * Tf.d[k]);
    acc[G_DIST] += -Av * ReS / dist;
    acc[G_PHI] += -R.m * Av * ImS;
    acc[G_COSI] += Q[4].re;
    acc[G_LAM] += Q[1].re - Av * ImS * pRd_l;
    acc[G_SINB] += (Q[2].re - Av * ImS * pRd_b) * R.inv_cosb;
    acc[G_PSI] += Q[3].re;
    acc[G_TREF] += Q[0].re - Av * ImS * (2. * PI * f + pRd_t);
}

// Forward pass and adjoint weights of one template (one wavefront): d_h, h_h
// as the forward kernel computes them, and gw[(b * 3 + c) * K + j] = the
// complex weight g of node j, channel c.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_grad_weights(
    double* __restrict__ d_h, double* __restrict__ h_h, cd* __restrict__ gw, const FusedRec* __restrict__ recs,
    const double* __restrict__ params, const double* __restrict__ t_start, const double* __restrict__ t_end, int B,
    int M, const double* __restrict__ freqs, const cd* __restrict__ h0, const cd* __restrict__ dc, int K,
    resp::OrbitTab orb, int TDItag, int rescaled, int tdi2, int jend, double initial_t_val)
{
    extern __shared__ cd rs[];  // residuals [3][K]
    const int b = blockIdx.x;
    Window win;
    win.t_ref = params[P_TREF * B + b];
    win.ts = t_start[b];
    win.te = t_end[b];
    win.dt_shift = win.t_ref - initial_t_val;
    residuals(rs, recs + (long)b * M, M, win, freqs, h0, K, orb, TDItag, rescaled, tdi2);
    __syncthreads();
    double sums[4];
    segment_sums(rs, freqs, dc, K, jend, sums);
    adjoint_weights(gw + (long)b * 3 * K, rs, freqs, h0, dc, K, jend);
    if (threadIdx.x == 0) {
        d_h[2 * b] = sums[0];
        d_h[2 * b + 1] = sums[1];
        h_h[2 * b] = sums[2];
        h_h[2 * b + 1] = sums[3];
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
    double acc[NGRAD];
    for (int p = 0; p < NGRAD; p++) acc[p] = 0.0;
    for (int j = threadIdx.x; j < K; j += 64) {
        const cd g[3] = {g_b[j], g_b[K + j], g_b[2 * K + j]};
        grad_point(G, orb, freqs[j], win, g, TDItag, rescaled, tdi2, dist, acc);
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
                                        int include_last_segment, void* stream)
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
    FusedRec* recs = static_cast<FusedRec*>(workspace_get(WS_FUSED, sizeof(FusedRec) * (size_t)nrec));
    GradRec* grecs = static_cast<GradRec*>(workspace_get(WS_GRAD, sizeof(GradRec) * (size_t)nrec));
    // adjoint weights (B, 3, K) complex, then partial gradients (B, M, 11)
    const size_t gw_bytes = sizeof(cd) * 3 * (size_t)K * numBinAll;
    char* scratch = static_cast<char*>(
        workspace_get(WS_GRAD_AUX, gw_bytes + sizeof(double) * NGRAD * (size_t)nrec));
    if (!recs || !grecs || !scratch) return ERR_ALLOC;
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
    hipLaunchKernelGGL(k_grad_setup, dim3(ceil_div((long)nrec, 64)), dim3(64), 0, side->stream, grecs, params,
                       numBinAll, ml, numModes, qnm, table_len, dspin);
    if (hipEventRecord(side->join, side->stream) != hipSuccess) return ERR_LAUNCH;
    hipLaunchKernelGGL(k_grad_setup_fwd, dim3(ceil_div((long)nrec, 64)), dim3(64), 0, st, recs, params, numBinAll, ml,
                       numModes, qnm, table_len, dspin);
    if (hipGetLastError() != hipSuccess) return ERR_LAUNCH;
    if (timed) (void)hipEventRecord(ev[1], st);
    const int jend = include_last_segment ? K : K - 1;
    hipLaunchKernelGGL(k_grad_weights, dim3(numBinAll), dim3(64), sizeof(cd) * 3 * (size_t)K, st, d_h, h_h, gw, recs,
                       params, t_start, t_end, numBinAll, numModes, freqs, reinterpret_cast<const cd*>(h0),
                       reinterpret_cast<const cd*>(dataConstants), K, orb, TDItag, rescaled, tdi2, jend,
                       initial_t_val);
    if (hipStreamWaitEvent(st, side->join, 0) != hipSuccess) return ERR_LAUNCH;
    hipLaunchKernelGGL(k_grad_points, dim3(nrec), dim3(64), 0, st, partial, grecs, gw, params, t_start, t_end,
                       numBinAll, numModes, freqs, K, orb, TDItag, rescaled, tdi2, initial_t_val);
    hipLaunchKernelGGL(k_grad_reduce, dim3(ceil_div((long)numBinAll * NGRAD, 256)), dim3(256), 0, st, grad, partial,
                       numBinAll, numModes);
    if (timed) (void)hipEventRecord(ev[2], st);
    return hipGetLastError() == hipSuccess ? OK : ERR_LAUNCH;
}
