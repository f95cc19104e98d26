// FP64 CPU restatement of the LISA fast frequency-domain TDI response.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Follows cutils/Response.cu of the reference (Marsat & Baker,
// arXiv:1806.10734, arXiv:2003.00357). The reference reads spacecraft
// positions and link unit vectors from lisaanalysistools' Orbits class
// (Detector.hpp, not part of the reference tree); here they come from tables
// in the same layout (response/fastfdresponse.py:107-134) with linear
// interpolation in time, the behaviour that class has after
// configure(linear_interp_setup=True) (response/fastfdresponse.py:174).

#include "oracle.h"

#include <cmath>

namespace orc {
namespace {

struct Orbit {
    double dt;
    int N;
    const double* n_arr;  // (N, 6, 3)
    const double* x_arr;  // (N, 3, 3)
    const int* links;     // (6,)
    double armlength;

    // Time window; times outside the table extrapolate from the first/last segment.
    int window(double t, double& rel) const
    {
        int w = int(std::floor(t / dt));
        if (w < 0) w = 0;
        if (w > N - 2) w = N - 2;
        rel = t - dt * w;
        return w;
    }

    void pos(double t, int sc, double out[3]) const
    {
        double rel;
        int w = window(t, rel);
        for (int k = 0; k < 3; k++) {
            double lo = x_arr[(w * 3 + (sc - 1)) * 3 + k];
            double hi = x_arr[((w + 1) * 3 + (sc - 1)) * 3 + k];
            out[k] = ((hi - lo) / dt) * rel + lo;
        }
    }

    void normal(double t, int link, double out[3]) const
    {
        int li = 0;
        for (int j = 0; j < 6; j++)
            if (links[j] == link) li = j;
        double rel;
        int w = window(t, rel);
        for (int k = 0; k < 3; k++) {
            double lo = n_arr[(w * 6 + li) * 3 + k];
            double hi = n_arr[((w + 1) * 6 + li) * 3 + k];
            out[k] = ((hi - lo) / dt) * rel + lo;
        }
    }
};

double dot3(const double* a, const double* b)
{
    double o = 0.0;
    for (int i = 0; i < 3; i++) o += a[i] * b[i];
    return o;
}

// n^T H n for complex H (Response.cu:23-39)
cplx nHn(const double* n, const cplx* H)
{
    cplx out(0.0, 0.0);
    for (int i = 0; i < 3; i++) {
        cplx t(0.0, 0.0);
        for (int j = 0; j < 3; j++) t += H[i * 3 + j] * n[j];
        out += n[i] * t;
    }
    return out;
}

double sinc(double x) { return x == 0.0 ? 1.0 : std::sin(x) / x; }

struct G6 {
    cplx G12, G21, G23, G32, G31, G13;
};

// single-link responses, Response.cu:54-194 (full response, orbital delay included)
G6 single_links(double t, double f, const cplx* H, const double* k, const double* p0, const Orbit& orb)
{
    const cplx I(0.0, 1.0);
    double p1[3], p2[3], p3[3], n[3];
    orb.pos(t, 1, p1);
    orb.pos(t, 2, p2);
    orb.pos(t, 3, p3);
    orb.normal(t, 12, n);
    double kn1 = dot3(k, n);
    cplx n1Hn1 = nHn(n, H);
    orb.normal(t, 23, n);
    double kn2 = dot3(k, n);
    cplx n2Hn2 = nHn(n, H);
    orb.normal(t, 31, n);
    double kn3 = dot3(k, n);
    cplx n3Hn3 = nHn(n, H);

    double s12[3], s23[3], s31[3];
    for (int i = 0; i < 3; i++) {
        s12[i] = p1[i] + p2[i];
        s23[i] = p2[i] + p3[i];
        s31[i] = p3[i] + p1[i];
    }
    double kp12 = dot3(k, s12), kp23 = dot3(k, s23), kp31 = dot3(k, s31), kp0 = dot3(k, p0);
    const double L = orb.armlength;
    cplx e0 = std::exp(I * 2. * PI * f / C_SI * kp0);
    double pre = PI * f * L / C_SI;
    cplx e12 = std::exp(I * pre * (1. + kp12 / L));
    cplx e23 = std::exp(I * pre * (1. + kp23 / L));
    cplx e31 = std::exp(I * pre * (1. + kp31 / L));
    cplx common = I * pre * e0;
    G6 g;
    g.G12 = common * n3Hn3 * sinc(pre * (1. - kn3)) * e12;
    g.G21 = common * n3Hn3 * sinc(pre * (1. + kn3)) * e12;
    g.G23 = common * n1Hn1 * sinc(pre * (1. - kn1)) * e23;
    g.G32 = common * n1Hn1 * sinc(pre * (1. + kn1)) * e23;
    g.G31 = common * n2Hn2 * sinc(pre * (1. - kn2)) * e31;
    g.G13 = common * n2Hn2 * sinc(pre * (1. + kn2)) * e31;
    return g;
}

// TDI combinations, Response.cu:197-261
void tdi_combination(const G6& g, double f, int TDItag, int rescaled, int tdi2, double L, cplx out[3])
{
    const cplx I(0.0, 1.0);
    double x = PI * f * L / C_SI;
    cplx z = std::exp(I * 2. * x);
    if (TDItag == 1) {
        cplx factor = rescaled ? cplx(1.0, 0.0) : 2. * I * std::sin(2. * x) * z;
        out[0] = factor * (g.G21 + z * g.G12 - g.G31 - z * g.G13);
        out[1] = factor * (g.G32 + z * g.G23 - g.G12 - z * g.G21);
        out[2] = factor * (g.G13 + z * g.G31 - g.G23 - z * g.G32);
        return;
    }
    cplx fAE(1.0, 0.0), fT(1.0, 0.0);
    if (!rescaled) {
        fAE = I * SQRT2 * std::sin(2. * x) * z;
        fT = 2. * SQRT2 * std::sin(2. * x) * std::sin(x) * std::exp(I * 3. * x);
    }
    cplx A = 0.5 * ((1. + z) * (g.G31 + g.G13) - g.G23 - z * g.G32 - g.G21 - z * g.G12);
    cplx E = 0.5 * INVSQRT3 * ((1. - z) * (g.G13 - g.G31) + (2. + z) * (g.G12 - g.G32) + (1. + 2. * z) * (g.G21 - g.G23));
    cplx T = INVSQRT6 * (g.G21 - g.G12 + g.G32 - g.G23 + g.G13 - g.G31);
    cplx f2 = tdi2 ? (-2. * I * std::sin(4. * x) * std::exp(I * 4. * x)) : cplx(1.0, 0.0);
    const cplx conv(2., 0.0);
    out[0] = f2 * conv * fAE * A;
    out[1] = f2 * conv * fAE * E;
    out[2] = f2 * conv * fT * T;
}

// spin-weight -2 spherical harmonics for the PhenomHM modes, Response.cu:383-424
cplx sYlm(int l, int m, double theta, double phi)
{
    double fac = 0.0;
    double c = std::cos(theta), ch = std::cos(theta / 2.0), sh = std::sin(theta / 2.0), st = std::sin(theta);
    if (l == 2 && m == -2) fac = std::sqrt(5.0 / (64.0 * PI)) * (1.0 - c) * (1.0 - c);
    else if (l == 2 && m == -1) fac = std::sqrt(5.0 / (16.0 * PI)) * st * (1.0 - c);
    else if (l == 2 && m == 1) fac = std::sqrt(5.0 / (16.0 * PI)) * st * (1.0 + c);
    else if (l == 2 && m == 2) fac = std::sqrt(5.0 / (64.0 * PI)) * (1.0 + c) * (1.0 + c);
    else if (l == 3 && m == -3) fac = std::sqrt(21.0 / (2.0 * PI)) * ch * std::pow(sh, 5.0);
    else if (l == 3 && m == -2) fac = std::sqrt(7.0 / (4.0 * PI)) * (2.0 + 3.0 * c) * std::pow(sh, 4.0);
    else if (l == 3 && m == 2) fac = std::sqrt(7.0 / PI) * std::pow(ch, 4.0) * (-2.0 + 3.0 * c) / 2.0;
    else if (l == 3 && m == 3) fac = -std::sqrt(21.0 / (2.0 * PI)) * std::pow(ch, 5.0) * sh;
    else if (l == 4 && m == -4) fac = 3.0 * std::sqrt(7.0 / PI) * std::pow(ch, 2.0) * std::pow(sh, 6.0);
    else if (l == 4 && m == -3) fac = 3.0 * std::sqrt(7.0 / (2.0 * PI)) * ch * (1.0 + 2.0 * c) * std::pow(sh, 5.0);
    else if (l == 4 && m == 3) fac = -3.0 * std::sqrt(7.0 / (2.0 * PI)) * std::pow(ch, 5.0) * (-1.0 + 2.0 * c) * sh;
    else if (l == 4 && m == 4) fac = 3.0 * std::sqrt(7.0 / PI) * std::pow(ch, 6.0) * std::pow(sh, 2.0);
    if (m == 0) return cplx(fac, 0.0);
    return fac * std::exp(cplx(0.0, 1.0) * cplx(m * phi, 0.0));
}

void matmul3(double* out, const double* a, const double* b)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            out[i * 3 + j] = 0.0;
            for (int k = 0; k < 3; k++) out[i * 3 + j] += a[i * 3 + k] * b[k * 3 + j];
        }
}

// Wave vector and polarisation tensors in the SSB frame, Response.cu:503-568:
// H+ = O1 diag(1, -1, 0) O1^-1, Hx = O1 offdiag(1, 1) O1^-1.
void polarization(double lam, double beta, double psi, double k[3], double Hp[9], double Hc[9])
{
    const double HSplus[9] = {1., 0., 0., 0., -1., 0., 0., 0., 0.};
    const double HScross[9] = {0., 1., 0., 1., 0., 0., 0., 0., 0.};
    double cl = std::cos(lam), sl = std::sin(lam);
    double cb = std::cos(beta), sb = std::sin(beta);
    double cp = std::cos(psi), sp = std::sin(psi);
    k[0] = -std::cos(beta) * std::cos(lam);
    k[1] = -std::cos(beta) * std::sin(lam);
    k[2] = -std::sin(beta);
    double O1[9] = {cp * sl - cl * sb * sp, -cl * cp * sb - sl * sp, -cb * cl,
                    -cl * cp - sb * sl * sp, -cp * sb * sl + cl * sp, -cb * sl,
                    cb * sp, cb * cp, -sb};
    double invO1[9] = {cp * sl - cl * sb * sp, -cl * cp - sb * sl * sp, cb * sp,
                       -cl * cp * sb - sl * sp, -cp * sb * sl + cl * sp, cb * cp,
                       -cb * cl, -cb * sl, -sb};
    double tmp[9];
    matmul3(tmp, HSplus, invO1);
    matmul3(Hp, O1, tmp);
    matmul3(tmp, HScross, invO1);
    matmul3(Hc, O1, tmp);
}

}  // namespace
}  // namespace orc

using namespace orc;

extern "C" {

// FastLISAResponse::LISA_response -> response -> responseCore -> response_modes,
// Response.cu:321-377, :468-608, :617-669, :671-746
void orc_lisa_response(double* buffer, const int* ells, const int* mms, const double* freqs,
                       const double* phi_ref, const double* inc, const double* lam,
                       const double* beta, const double* psi, int TDItag, int rescaled,
                       int tdi2, int numModes, int length, int numBinAll, int includesAmps,
                       double orbit_dt, int orbit_N, const double* n_arr, const double* x_arr,
                       const int* links, double armlength)
{
    const long nper = (long)numBinAll * numModes * length;
    double* phases = buffer + includesAmps * nper;
    const double* tf = buffer + (includesAmps + 1) * nper;
    double* resp = buffer + (includesAmps + 2) * nper;
    Orbit orb{orbit_dt, orbit_N, n_arr, x_arr, links, armlength};
    const cplx I(0.0, 1.0);

#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < numBinAll; b++) {
        double k[3], Hp[9], Hc[9];
        polarization(lam[b], beta[b], psi[b], k, Hp, Hc);

        for (int mi = 0; mi < numModes; mi++) {
            int l = ells[mi], m = mms[mi];
            cplx Ylm = sYlm(l, m, inc[b], phi_ref[b]);
            cplx Yl_m = std::pow(-1.0, l) * std::conj(sYlm(l, -1 * m, inc[b], phi_ref[b]));
            cplx Yp = 1. / 2 * (Ylm + Yl_m);
            cplx Yc = 1. / 2. * I * (Ylm - Yl_m);
            cplx H[9];
            for (int q = 0; q < 9; q++) H[q] = (Yp * Hp[q] + Yc * Hc[q]);

            for (int i = 0; i < length; i++) {
                long idx = ((long)b * numModes + mi) * length + i;
                double f = freqs[idx];
                double t = tf[idx];
                // guiding-centre position and orbital delay, Response.cu:264-314
                double alpha = Omega0 * t;
                double p0[3] = {AU_SI * std::cos(alpha), AU_SI * std::sin(alpha), 0. * t};
                double kR = dot3(k, p0);
                double phaseRdelay = 2. * PI / C_SI * f * kR;
                G6 g = single_links(t, f, H, k, p0, orb);
                cplx rot = std::exp(cplx(0.0, -1.0) * phaseRdelay);
                G6 T;
                T.G12 = g.G12 * rot;
                T.G21 = g.G21 * rot;
                T.G23 = g.G23 * rot;
                T.G32 = g.G32 * rot;
                T.G31 = g.G31 * rot;
                T.G13 = g.G13 * rot;
                cplx out[3];
                tdi_combination(T, f, TDItag, rescaled, tdi2, armlength, out);
                resp[0 * nper + idx] = out[0].real();
                resp[1 * nper + idx] = out[0].imag();
                resp[2 * nper + idx] = out[1].real();
                resp[3 * nper + idx] = out[1].imag();
                resp[4 * nper + idx] = out[2].real();
                resp[5 * nper + idx] = out[2].imag();
                phases[idx] += phaseRdelay;
            }
        }
    }
}

// Spin-weight -2 spherical harmonic of the response (Response.cu:383-424);
// out = (re, im). Exported for the oracle's own pins.
void orc_sylm(int l, int m, double theta, double phi, double* out)
{
    const cplx y = sYlm(l, m, theta, phi);
    out[0] = y.real();
    out[1] = y.imag();
}

// Wave vector k (3) and polarisation tensors H+ (9), Hx (9), row-major
// (Response.cu:503-568). Exported for the oracle's own pins.
void orc_polarization(double lam, double beta, double psi, double* k, double* Hp, double* Hc)
{
    polarization(lam, beta, psi, k, Hp, Hc);
}

}  // extern "C"

