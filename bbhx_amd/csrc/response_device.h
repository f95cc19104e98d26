// LISA fast frequency-domain TDI response on gfx950.
//
// Physics follows the reference's cutils/Response.cu (Marsat & Baker,
// arXiv:1806.10734; arXiv:2003.00357). Spacecraft positions and link unit
// vectors come from orbit tables in the lisaanalysistools Orbits layout
// (x: (N, 3, 3), n: (N, 6, 3), uniform step dt), linearly interpolated in time.
//
// Per (binary, mode) resp_setup builds the polarisation basis and the
// spherical-harmonic factors once; per frequency point resp_eval returns the
// three complex transfer functions and the orbital Doppler phase.
//
// Exact rewrites relative to the reference's operation order (they change
// results only at the level of FP64 rounding):
//  * H = Y+ H+ + Yx Hx with H+ = u u^T - v v^T, Hx = u v^T + v u^T (u, v the
//    first two columns of the source-frame rotation, Response.cu:540-567), so
//    n^T H n = Y+ ((u.n)^2 - (v.n)^2) + 2 Yx (u.n)(v.n): no 3x3 complex matrix.
//  * linear interpolation commutes with dot products: k.p_i, k.n, u.n, v.n are
//    interpolated directly from their values at the two table nodes.
//  * the guiding-centre factor exp(i 2 pi f k.p0 / c) (Response.cu:166) and the
//    rotation exp(-i phaseRdelay) (Response.cu:304-309) cancel identically.
//  * divisions by L and by x (1 +- kn) become products with reciprocals.
//  * sin(x(1 +- kn)) from sincos(x) and sincos(x kn) by angle addition; every
//    function of x = pi f L / c from one sincos(x) by multiple-angle identities.
//  * arguments of e^{i theta} with large |theta| are reduced modulo 2 pi in
//    two steps (Cody-Waite) before sincos.
#pragma once

#include "common.h"
#include "dual.h"

namespace bbhx {
namespace resp {

struct RespRec {
    double u[3], v[3];  // polarisation basis in the SSB frame
    double k[3];        // wave propagation vector
    cd Yp, Yc;          // (Ylm + (-1)^l Yl-m*)/2 and i (Ylm - (-1)^l Yl-m*)/2
};

// Packed orbit rows (built once per orbit table, bbhx_hip_response_add_orbits):
// row w holds the 18 values P_w = (x of spacecraft 1, 2, 3; unit vectors of
// links 12, 23, 31) at table time w dt, then the 18 differences P_{w+1} - P_w
// (zero in the last row), so that one contiguous 288-B row gives the linear
// interpolation at any t in [w dt, (w + 1) dt).
constexpr int ORB_VALS = 18;
constexpr int ORB_ROW = 2 * ORB_VALS;

struct OrbitTab {
    const double* x;  // (N, 3, 3)
    const double* n;  // (N, 6, 3)
    const double* p;  // (N, ORB_ROW) packed rows
    double dt, inv_dt;
    int N;
    int li12, li23, li31;  // row of links 12, 23, 31 in n
    double L;              // armlength [m]
    double inv_L;          // 1 / L
};

// Real factor of the spin-weight -2 spherical harmonic -2Y_lm(theta, phi) =
// fac(theta) e^{i m phi} for the PhenomHM (l, m) (Response.cu:383-424).
// Written in the half-angle cosine ch and sine sh of theta (1 + cos = 2 ch^2,
// 1 - cos = 2 sh^2, sin = 2 sh ch, cos = ch^2 - sh^2) with integer powers as
// products: two trigonometric calls and no pow.
template <class S>
__device__ inline S ylm_fac(int l, int m, S theta)
{
    S fac = 0.0;
    const S ch = cos(theta / 2.0), sh = sin(theta / 2.0);
    const S ch2 = ch * ch, sh2 = sh * sh;
    const S c = ch2 - sh2;
    if (l == 2 && m == -2) fac = (4.0 * sqrt(5.0 / (64.0 * PI))) * (sh2 * sh2);
    else if (l == 2 && m == -1) fac = (4.0 * sqrt(5.0 / (16.0 * PI))) * (sh2 * sh * ch);
    else if (l == 2 && m == 1) fac = (4.0 * sqrt(5.0 / (16.0 * PI))) * (ch2 * ch * sh);
    else if (l == 2 && m == 2) fac = (4.0 * sqrt(5.0 / (64.0 * PI))) * (ch2 * ch2);
    else if (l == 3 && m == -3) fac = sqrt(21.0 / (2.0 * PI)) * ch * (sh2 * sh2 * sh);
    else if (l == 3 && m == -2) fac = sqrt(7.0 / (4.0 * PI)) * (2.0 + 3.0 * c) * (sh2 * sh2);
    else if (l == 3 && m == 2) fac = sqrt(7.0 / PI) * (ch2 * ch2) * (-2.0 + 3.0 * c) / 2.0;
    else if (l == 3 && m == 3) fac = -sqrt(21.0 / (2.0 * PI)) * (ch2 * ch2 * ch) * sh;
    else if (l == 4 && m == -4) fac = 3.0 * sqrt(7.0 / PI) * ch2 * (sh2 * sh2 * sh2);
    else if (l == 4 && m == -3) fac = 3.0 * sqrt(7.0 / (2.0 * PI)) * ch * (1.0 + 2.0 * c) * (sh2 * sh2 * sh);
    else if (l == 4 && m == 3) fac = -3.0 * sqrt(7.0 / (2.0 * PI)) * (ch2 * ch2 * ch) * (-1.0 + 2.0 * c) * sh;
    else if (l == 4 && m == 4) fac = 3.0 * sqrt(7.0 / PI) * (ch2 * ch2 * ch2) * sh2;
    return fac;
}

__device__ inline cd sYlm(int l, int m, double theta, double phi)
{
    const double fac = ylm_fac(l, m, theta);
    if (m == 0) return mk(fac, 0.0);
    return fac * cexpi(m * phi);
}

// Polarisation basis, wave vector and Ylm factors of one (binary, mode)
// (Response.cu:501-601).
__device__ inline void resp_setup(RespRec& r, int l, int m, double phi_ref, double inc, double lam, double beta,
                                  double psi)
{
    double cl = cos(lam), sl = sin(lam), cb = cos(beta), sb = sin(beta), cp = cos(psi), sp = sin(psi);
    r.k[0] = -cb * cl;
    r.k[1] = -cb * sl;
    r.k[2] = -sb;
    // first two columns of O1
    r.u[0] = cp * sl - cl * sb * sp;
    r.u[1] = -cl * cp - sb * sl * sp;
    r.u[2] = cb * sp;
    r.v[0] = -cl * cp * sb - sl * sp;
    r.v[1] = -cp * sb * sl + cl * sp;
    r.v[2] = cb * cp;
    cd Ylm = sYlm(l, m, inc, phi_ref);
    cd Yl_m = conj(sYlm(l, -m, inc, phi_ref));
    if (l & 1) Yl_m = mk(-Yl_m.re, -Yl_m.im);
    r.Yp = 0.5 * (Ylm + Yl_m);
    r.Yc = mul_i(0.5 * (Ylm - Yl_m));
}

__device__ inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Orbit quantities at time t that the response needs:
//   kp[i] = k . p_i (i = spacecraft 1..3), and for links 12, 23, 31:
//   kn[j] = k . n_j, un[j] = u . n_j, vn[j] = v . n_j
// from the positions and link vectors linearly interpolated in time (the
// table's lisatools linear interpolation), read as one packed row.
__device__ inline void orbit_eval(const OrbitTab& o, const RespRec& r, double t, double kp[3], double kn[3],
                                  double un[3], double vn[3])
{
    int w = int(floor(t * o.inv_dt));
    w = w < 0 ? 0 : (w > o.N - 2 ? o.N - 2 : w);
    const double a = (t - o.dt * w) * o.inv_dt;  // interpolation weight in [0, 1) inside the table
    const double2* row = reinterpret_cast<const double2*>(o.p + (long)w * ORB_ROW);
    double v[ORB_VALS];
    for (int q = 0; q < ORB_VALS / 2; q++) {
        const double2 lo = row[q], d = row[ORB_VALS / 2 + q];
        v[2 * q] = fma(d.x, a, lo.x);
        v[2 * q + 1] = fma(d.y, a, lo.y);
    }
    for (int i = 0; i < 3; i++) kp[i] = dot3(r.k, v + 3 * i);
    for (int j = 0; j < 3; j++) {
        const double* nj = v + 9 + 3 * j;
        kn[j] = dot3(r.k, nj);
        un[j] = dot3(r.u, nj);
        vn[j] = dot3(r.v, nj);
    }
}

// Quantities of one frequency that every (binary, mode) shares: x = pi f L / c
// (the reference's 'prefactor') and its sine and cosine.
struct NodeResp {
    double x, sx, cx;
};

__device__ inline NodeResp resp_node(const OrbitTab& o, double f)
{
    NodeResp n;
    n.x = PI * f * o.L / C_SI;
    sincos_red(n.x, &n.sx, &n.cx);
    return n;
}

// The six single-link terms of one (binary, mode) at frequency f and time t,
// G[0..5] = G12, G21, G23, G32, G31, G13 (Response.cu:54-194), and the
// orbital Doppler phase in 'phaseRdelay' (Response.cu:278-309); e^{i theta}
// from the table 'tab' (cexpi_tab).
// As resp_links, with the Y+ and Yx factors given (Yp, Yc in place of r.Yp,
// r.Yc): the link terms are linear in them, so several modes that share f and t
// give their weighted sum from one call.
template <class E>
__device__ inline void resp_links_y(const RespRec& r, cd Yp, cd Yc, const OrbitTab& o, const NodeResp& nr, double f,
                                    double t, cd G[6], double& phaseRdelay, const E* __restrict__ tab)
{
    double kp[3], kn[3], un[3], vn[3];
    orbit_eval(o, r, t, kp, kn, un, vn);
    const cd eo = cexpi_tab(OMEGA0 * t, tab);
    const double kR = r.k[0] * (AU_SI * eo.re) + r.k[1] * (AU_SI * eo.im);
    phaseRdelay = 2. * PI / C_SI * f * kR;

    const double x = nr.x, sx = nr.sx, cx = nr.cx;
    // link j = 0, 1, 2 is n1 (12), n2 (23), n3 (31); n_j^T H n_j
    cd nHn[3];
    double sp[3], sm[3];  // sm: sinc(x (1 - kn)), sp: sinc(x (1 + kn))
    for (int j = 0; j < 3; j++) {
        nHn[j] = Yp * (un[j] * un[j] - vn[j] * vn[j]) + Yc * (2.0 * un[j] * vn[j]);
        const cd e = cexpi_tab(x * kn[j], tab);
        const double s = e.im, c = e.re;
        // 1/am and 1/ap from one reciprocal of their product
        const double am = x * (1. - kn[j]), ap = x * (1. + kn[j]);
        const double prod = am * ap;
        if (prod != 0.0) {
            const double iprod = rcp_pos(prod);
            sm[j] = (sx * c - cx * s) * (ap * iprod);
            sp[j] = (sx * c + cx * s) * (am * iprod);
        } else {  // f = 0 or a link along k: sinc(0) = 1
            sm[j] = am == 0.0 ? 1.0 : (sx * c - cx * s) / am;
            sp[j] = ap == 0.0 ? 1.0 : (sx * c + cx * s) / ap;
        }
    }
    // exp(i x (1 + k.(p_i + p_j)/L)) for the pairs 12, 23, 31 (Response.cu:171-173)
    const cd e12 = cexpi_tab(x * (1. + (kp[0] + kp[1]) * o.inv_L), tab);
    const cd e23 = cexpi_tab(x * (1. + (kp[1] + kp[2]) * o.inv_L), tab);
    const cd e31 = cexpi_tab(x * (1. + (kp[2] + kp[0]) * o.inv_L), tab);
    const cd g3 = mul_i(x * nHn[2]), g1 = mul_i(x * nHn[0]), g2 = mul_i(x * nHn[1]);
    G[0] = (g3 * sm[2]) * e12;
    G[1] = (g3 * sp[2]) * e12;
    G[2] = (g1 * sm[0]) * e23;
    G[3] = (g1 * sp[0]) * e23;
    G[4] = (g2 * sm[1]) * e31;
    G[5] = (g2 * sp[1]) * e31;
}

template <class E>
__device__ inline void resp_links(const RespRec& r, const OrbitTab& o, const NodeResp& nr, double f, double t,
                                  cd G[6], double& phaseRdelay, const E* __restrict__ tab)
{
    resp_links_y(r, r.Yp, r.Yc, o, nr, f, t, G, phaseRdelay, tab);
}

// TDI channels 1..3 from link terms G[0..5] (order of resp_links) at the
// frequency of 'nr' (Response.cu:197-261). Linear in G, so it may be applied
// to a weighted sum of several modes' link terms at one frequency.
__device__ inline void tdi_combine(const cd G[6], const NodeResp& nr, int TDItag, int rescaled, int tdi2, cd out[3])
{
    const cd G12 = G[0], G21 = G[1], G23 = G[2], G32 = G[3], G31 = G[4], G13 = G[5];
    const double sx = nr.sx, cx = nr.cx;
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

// Transfer functions of TDI channels 1..3 at frequency f and time t; returns
// the orbital Doppler phase in 'phaseRdelay' (Response.cu:54-314).
__device__ inline void resp_eval(const RespRec& r, const OrbitTab& o, double f, double t, int TDItag, int rescaled,
                                 int tdi2, cd out[3], double& phaseRdelay)
{
    const NodeResp nr = resp_node(o, f);
    cd G[6];
    resp_links(r, o, nr, f, t, G, phaseRdelay, ETAB);
    tdi_combine(G, nr, TDItag, rescaled, tdi2, out);
}

}  // namespace resp
}  // namespace bbhx
