// LISA fast frequency-domain TDI response on gfx950.
//
// Physics follows the reference's cutils/Response.cu (Marsat & Baker,
// arXiv:1806.10734; arXiv:2003.00357). Spacecraft positions and link unit
// vectors come from orbit tables in the lisaanalysistools Orbits layout
// (x: (N, 3, 3), n: (N, 6, 3), uniform step dt), linearly interpolated in time.
//
// Per (binary, mode) the polarisation matrix H and the wave vector are built
// once (resp_setup); per frequency point resp_eval returns the three complex
// transfer functions and the orbital Doppler phase.
//
// Exact algebraic simplifications relative to the reference's operation order:
//  * the guiding-centre factor exp(i 2 pi f k.p0 / c) (Response.cu:166) and the
//    rotation exp(-i phaseRdelay) (Response.cu:304-309) cancel identically; both
//    are skipped.
//  * all functions of x = pi f L / c (z = e^{2ix}, sin 2x, e^{3ix}, sin 4x,
//    e^{4ix}) are built from one sincos(x) by multiple-angle identities.
// Both change results only at the level of FP64 rounding.
#pragma once

#include "common.h"

namespace bbhx {
namespace resp {

struct RespRec {
    cd H[9];      // Yfactor+ H+ + Yfactorx Hx (Response.cu:580-601)
    double k[3];  // wave propagation vector
};

struct OrbitTab {
    const double* x;  // (N, 3, 3)
    const double* n;  // (N, 6, 3)
    double dt, inv_dt;
    int N;
    int li12, li23, li31;  // row of links 12, 23, 31 in n
    double L;              // armlength [m]
};

// spin-weight -2 spherical harmonic for the PhenomHM (l, m) (Response.cu:383-424)
__device__ inline cd sYlm(int l, int m, double theta, double phi)
{
    double fac = 0.0;
    double c = cos(theta), ch = cos(theta / 2.0), sh = sin(theta / 2.0), st = sin(theta);
    if (l == 2 && m == -2) fac = sqrt(5.0 / (64.0 * PI)) * (1.0 - c) * (1.0 - c);
    else if (l == 2 && m == -1) fac = sqrt(5.0 / (16.0 * PI)) * st * (1.0 - c);
    else if (l == 2 && m == 1) fac = sqrt(5.0 / (16.0 * PI)) * st * (1.0 + c);
    else if (l == 2 && m == 2) fac = sqrt(5.0 / (64.0 * PI)) * (1.0 + c) * (1.0 + c);
    else if (l == 3 && m == -3) fac = sqrt(21.0 / (2.0 * PI)) * ch * pow(sh, 5.0);
    else if (l == 3 && m == -2) fac = sqrt(7.0 / (4.0 * PI)) * (2.0 + 3.0 * c) * pow(sh, 4.0);
    else if (l == 3 && m == 2) fac = sqrt(7.0 / PI) * pow(ch, 4.0) * (-2.0 + 3.0 * c) / 2.0;
    else if (l == 3 && m == 3) fac = -sqrt(21.0 / (2.0 * PI)) * pow(ch, 5.0) * sh;
    else if (l == 4 && m == -4) fac = 3.0 * sqrt(7.0 / PI) * pow(ch, 2.0) * pow(sh, 6.0);
    else if (l == 4 && m == -3) fac = 3.0 * sqrt(7.0 / (2.0 * PI)) * ch * (1.0 + 2.0 * c) * pow(sh, 5.0);
    else if (l == 4 && m == 3) fac = -3.0 * sqrt(7.0 / (2.0 * PI)) * pow(ch, 5.0) * (-1.0 + 2.0 * c) * sh;
    else if (l == 4 && m == 4) fac = 3.0 * sqrt(7.0 / PI) * pow(ch, 6.0) * pow(sh, 2.0);
    if (m == 0) return mk(fac, 0.0);
    return fac * cexpi(m * phi);
}

// H matrix and wave vector for one (binary, mode) (Response.cu:501-601)
__device__ inline void resp_setup(RespRec& r, int l, int m, double phi_ref, double inc, double lam, double beta,
                                  double psi)
{
    double cl = cos(lam), sl = sin(lam), cb = cos(beta), sb = sin(beta), cp = cos(psi), sp = sin(psi);
    r.k[0] = -cb * cl;
    r.k[1] = -cb * sl;
    r.k[2] = -sb;
    // O1 and its inverse (transpose)
    const double O1[9] = {cp * sl - cl * sb * sp, -cl * cp * sb - sl * sp, -cb * cl,
                          -cl * cp - sb * sl * sp, -cp * sb * sl + cl * sp, -cb * sl,
                          cb * sp, cb * cp, -sb};
    // H+ = O1 diag(1,-1,0) O1^T, Hx = O1 offdiag(1,1) O1^T
    double Hp[9], Hc[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double u_i = O1[i * 3 + 0], v_i = O1[i * 3 + 1];
            double u_j = O1[j * 3 + 0], v_j = O1[j * 3 + 1];
            Hp[i * 3 + j] = u_i * u_j - v_i * v_j;
            Hc[i * 3 + j] = u_i * v_j + v_i * u_j;
        }
    cd Ylm = sYlm(l, m, inc, phi_ref);
    cd Yl_m = conj(sYlm(l, -m, inc, phi_ref));
    if (l & 1) Yl_m = mk(-Yl_m.re, -Yl_m.im);
    cd Yp = 0.5 * (Ylm + Yl_m);
    cd Yc = mul_i(0.5 * (Ylm - Yl_m));
    for (int q = 0; q < 9; q++) r.H[q] = Yp * Hp[q] + Yc * Hc[q];
}

__device__ inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__device__ inline cd nHn(const double* n, const cd* H)
{
    cd out = mk(0.0, 0.0);
    for (int i = 0; i < 3; i++) {
        cd t = H[i * 3 + 0] * n[0] + H[i * 3 + 1] * n[1] + H[i * 3 + 2] * n[2];
        out += n[i] * t;
    }
    return out;
}

__device__ inline double sinc(double x) { return x == 0.0 ? 1.0 : sin(x) / x; }

// Interpolated orbit quantities at time t: p1+p2, p2+p3, p3+p1 and n12, n23, n31.
__device__ inline void orbit_eval(const OrbitTab& o, double t, double s12[3], double s23[3], double s31[3],
                                  double n1[3], double n2[3], double n3[3])
{
    int w = int(floor(t * o.inv_dt));
    w = w < 0 ? 0 : (w > o.N - 2 ? o.N - 2 : w);
    const double rel = t - o.dt * w;
    const double* x0 = o.x + (long)w * 9;
    const double* x1 = x0 + 9;
    double p[9];
    for (int q = 0; q < 9; q++) p[q] = ((x1[q] - x0[q]) * o.inv_dt) * rel + x0[q];
    for (int k = 0; k < 3; k++) {
        s12[k] = p[0 * 3 + k] + p[1 * 3 + k];
        s23[k] = p[1 * 3 + k] + p[2 * 3 + k];
        s31[k] = p[2 * 3 + k] + p[0 * 3 + k];
    }
    const double* n0 = o.n + (long)w * 18;
    const double* nn = n0 + 18;
    for (int k = 0; k < 3; k++) {
        n1[k] = ((nn[o.li12 * 3 + k] - n0[o.li12 * 3 + k]) * o.inv_dt) * rel + n0[o.li12 * 3 + k];
        n2[k] = ((nn[o.li23 * 3 + k] - n0[o.li23 * 3 + k]) * o.inv_dt) * rel + n0[o.li23 * 3 + k];
        n3[k] = ((nn[o.li31 * 3 + k] - n0[o.li31 * 3 + k]) * o.inv_dt) * rel + n0[o.li31 * 3 + k];
    }
}

// Transfer functions of TDI channels 1..3 at frequency f and time t; returns
// the orbital Doppler phase in 'phaseRdelay' (Response.cu:54-314).
__device__ inline void resp_eval(const RespRec& r, const OrbitTab& o, double f, double t, int TDItag, int rescaled,
                                 int tdi2, cd out[3], double& phaseRdelay)
{
    double s12[3], s23[3], s31[3], n1[3], n2[3], n3[3];
    orbit_eval(o, t, s12, s23, s31, n1, n2, n3);
    const double* k = r.k;
    double sa, ca;
    sincos(OMEGA0 * t, &sa, &ca);
    const double kR = k[0] * (AU_SI * ca) + k[1] * (AU_SI * sa);
    phaseRdelay = 2. * PI / C_SI * f * kR;

    const double kn1 = dot3(k, n1), kn2 = dot3(k, n2), kn3 = dot3(k, n3);
    const cd n1Hn1 = nHn(n1, r.H), n2Hn2 = nHn(n2, r.H), n3Hn3 = nHn(n3, r.H);
    const double L = o.L;
    const double pre = PI * f * L / C_SI;
    // exp(i pre (1 + k.(pi+pj)/L)) (Response.cu:171-173)
    const cd e12 = cexpi(pre * (1. + dot3(k, s12) / L));
    const cd e23 = cexpi(pre * (1. + dot3(k, s23) / L));
    const cd e31 = cexpi(pre * (1. + dot3(k, s31) / L));
    // common factor i * pre
    const cd g3 = mul_i(pre * n3Hn3), g1 = mul_i(pre * n1Hn1), g2 = mul_i(pre * n2Hn2);
    const cd G12 = (g3 * sinc(pre * (1. - kn3))) * e12;
    const cd G21 = (g3 * sinc(pre * (1. + kn3))) * e12;
    const cd G23 = (g1 * sinc(pre * (1. - kn1))) * e23;
    const cd G32 = (g1 * sinc(pre * (1. + kn1))) * e23;
    const cd G31 = (g2 * sinc(pre * (1. - kn2))) * e31;
    const cd G13 = (g2 * sinc(pre * (1. + kn2))) * e31;

    // TDI combinations (Response.cu:197-261)
    const double x = pre;
    double sx, cx;
    sincos(x, &sx, &cx);
    const double s2x = 2.0 * sx * cx, c2x = cx * cx - sx * sx;
    const cd z = mk(c2x, s2x);
    if (TDItag == 1) {
        cd fac = rescaled ? mk(1.0, 0.0) : mul_i((2. * s2x) * z);
        out[0] = fac * (G21 + z * G12 - G31 - z * G13);
        out[1] = fac * (G32 + z * G23 - G12 - z * G21);
        out[2] = fac * (G13 + z * G31 - G23 - z * G32);
        return;
    }
    cd fAE = mk(1.0, 0.0), fT = mk(1.0, 0.0);
    if (!rescaled) {
        fAE = mul_i((SQRT2 * s2x) * z);
        const double c3x = c2x * cx - s2x * sx, s3x = s2x * cx + c2x * sx;
        fT = (2. * SQRT2 * s2x * sx) * mk(c3x, s3x);
    }
    const cd one = mk(1.0, 0.0);
    const cd A = 0.5 * ((one + z) * (G31 + G13) - G23 - z * G32 - G21 - z * G12);
    const cd E = (0.5 * INVSQRT3) * ((one - z) * (G13 - G31) + (mk(2.0, 0.0) + z) * (G12 - G32) +
                                    (one + 2. * z) * (G21 - G23));
    const cd T = INVSQRT6 * (G21 - G12 + G32 - G23 + G13 - G31);
    cd f2 = mk(1.0, 0.0);
    if (tdi2) {
        const double s4x = 2.0 * s2x * c2x, c4x = c2x * c2x - s2x * s2x;
        f2 = mul_i((-2. * s4x) * mk(c4x, s4x));
    }
    const cd fa = 2.0 * (f2 * fAE), ft = 2.0 * (f2 * fT);
    out[0] = fa * A;
    out[1] = fa * E;
    out[2] = ft * T;
}

}  // namespace resp
}  // namespace bbhx
