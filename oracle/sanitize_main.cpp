// Memory and undefined-behaviour check of the CPU oracle (TEST INFRASTRUCTURE).
//
// Built by `make -C oracle sanitize` with AddressSanitizer and
// UndefinedBehaviorSanitizer and run there: every oracle entry point on a small
// synthetic problem (2 binaries, 6 modes, 64 sparse points, an equal-arm
// constellation on a circular orbit tabulated at 200 times, 4096 dense bins).
// The sanitizers abort on an out-of-bounds access, a use after free or
// undefined arithmetic; the driver itself checks that every output is finite.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "oracle.h"

using namespace orc;

static int check_finite(const char* what, const std::vector<double>& v)
{
    for (double x : v)
        if (!std::isfinite(x)) {
            std::printf("non-finite value in %s\n", what);
            return 1;
        }
    return 0;
}

int main()
{
    const int B = 2, M = 6, L = 64, N = 4096;
    const int ells[M] = {2, 3, 4, 2, 3, 4}, mms[M] = {2, 3, 4, 1, 2, 3};
    const double m1[B] = {8e5, 3e6}, m2[B] = {3e5, 1e6};
    const double c1[B] = {0.6, -0.2}, c2[B] = {0.1, 0.4};
    double m1_SI[B], m2_SI[B];
    for (int b = 0; b < B; b++) {
        m1_SI[b] = m1[b] * MSUN_SI;
        m2_SI[b] = m2[b] * MSUN_SI;
    }
    const double dist[B] = {1e26, 3e26}, fref[B] = {0.0, 0.0};
    const double phi[B] = {0.3, 2.0}, inc[B] = {1.0, 2.2}, lam[B] = {0.7, 4.0}, beta[B] = {-0.3, 0.9},
                 psi[B] = {0.4, 2.5};
    int fails = 0;

    // ringdown frequencies: a smooth stand-in QNM table (the physics is not checked here)
    const int T = 1001;
    std::vector<double> tab(8 * T);
    for (int k = 0; k < T; k++) {
        const double a = -1.0 + 0.002 * k;
        tab[k] = 0.05 + 0.02 * a;
        tab[4 * T + k] = 0.012 + 0.001 * a;
    }
    std::vector<double> frd(B), fdm(B), hrd(B * M), hdm(B * M);
    orc_get_phenomd_ringdown_frequencies(frd.data(), fdm.data(), m1, m2, c1, c2, B, &tab[0], &tab[T], &tab[2 * T],
                                         &tab[3 * T], &tab[4 * T], &tab[5 * T], &tab[6 * T], &tab[7 * T], 0.002);
    orc_get_phenomhm_ringdown_frequencies(hrd.data(), hdm.data(), m1, m2, c1, c2, ells, mms, M, B);
    std::vector<double> Mf_RD(B * (M + 1)), Mf_DM(B * (M + 1));
    for (int b = 0; b < B; b++) {
        for (int m = 0; m < M; m++) {
            Mf_RD[b * (M + 1) + m] = hrd[b * M + m];
            Mf_DM[b * (M + 1) + m] = hdm[b * M + m];
        }
        Mf_RD[b * (M + 1) + M] = frd[b];
        Mf_DM[b * (M + 1) + M] = fdm[b];
    }
    fails += check_finite("ringdown", hrd) + check_finite("phenomd ringdown", frd);

    // PhenomHM on log-spaced per-mode grids
    std::vector<double> freqs(B * M * L), buf(9 * B * M * L, 0.0);
    for (int b = 0; b < B; b++) {
        const double Msec = (m1[b] + m2[b]) * MTSUN_SI;
        for (int m = 0; m < M; m++)
            for (int i = 0; i < L; i++) {
                const double Mf = 1e-4 * std::pow(0.6 / 1e-4, i / double(L - 1));
                freqs[(b * M + m) * L + i] = Mf / Msec * mms[m] / 2.0;
            }
    }
    orc_waveform_amp_phase(buf.data(), ells, mms, freqs.data(), m1_SI, m2_SI, c1, c2, dist, fref, M, L, B,
                           Mf_RD.data(), Mf_DM.data(), 0);
    // keep t(f) inside the orbit table
    const long nper = (long)B * M * L;
    for (long i = 0; i < nper; i++) buf[2 * nper + i] = 1e6 + std::fmod(std::fabs(buf[2 * nper + i]), 1e7);

    // equal-arm constellation on a circular 1-AU orbit, 200 samples
    const int NO = 200;
    const double dt = 1e5, Lr = 2.5e9;
    std::vector<double> x(NO * 9), n(NO * 18);
    const int links[6] = {12, 23, 31, 13, 32, 21};
    for (int w = 0; w < NO; w++) {
        const double a = Omega0 * w * dt;
        double p[3][3];
        for (int s = 0; s < 3; s++) {
            const double ph = a + 2.0 * PI * s / 3.0;
            p[s][0] = AU_SI * std::cos(a) + Lr / std::sqrt(3.0) * std::cos(ph) * 0.5;
            p[s][1] = AU_SI * std::sin(a) + Lr / std::sqrt(3.0) * std::sin(ph) * 0.5;
            p[s][2] = Lr / std::sqrt(3.0) * std::cos(ph) * std::sqrt(3.0) / 2.0;
            for (int c = 0; c < 3; c++) x[w * 9 + s * 3 + c] = p[s][c];
        }
        for (int j = 0; j < 6; j++) {
            const int r = links[j] / 10 - 1, e = links[j] % 10 - 1;
            double d[3], nn = 0.0;
            for (int c = 0; c < 3; c++) {
                d[c] = p[r][c] - p[e][c];
                nn += d[c] * d[c];
            }
            for (int c = 0; c < 3; c++) n[w * 18 + j * 3 + c] = d[c] / std::sqrt(nn);
        }
    }
    orc_lisa_response(buf.data(), ells, mms, freqs.data(), phi, inc, lam, beta, psi, 2, 0, 1, M, L, B, 1, dt, NO,
                      n.data(), x.data(), links, Lr);
    fails += check_finite("response", buf);

    // spline of the nine quantities and the dense template build
    std::vector<double> Bc(9 * nper), sc1(9 * nper), sc2(9 * nper), sc3(9 * nper);
    orc_interpolate(freqs.data(), buf.data(), Bc.data(), sc1.data(), sc2.data(), sc3.data(), L, 9, M, B);
    fails += check_finite("spline", sc3);
    std::vector<double> df(N);
    for (int k = 0; k < N; k++) df[k] = 1e-5 + 2e-6 * k;
    std::vector<int> st(B), len(B), inds;
    std::vector<long> toff(B), ioff(B);
    long tt = 0, it = 0;
    for (int b = 0; b < B; b++) {
        const double lo = freqs[(b * M) * L], hi = freqs[(b * M) * L + L - 1];
        int s = 0;
        while (s < N && df[s] <= lo) s++;
        int e = s;
        while (e < N && df[e] <= hi) e++;
        st[b] = s;
        len[b] = e - s;
        toff[b] = tt;  // complex elements
        ioff[b] = it;
        tt += 3L * len[b];
        it += (long)M * len[b];
        for (int m = 0; m < M; m++)
            for (int k = s; k < e; k++) {
                const double* xm = &freqs[(b * M + m) * L];
                int i = -1;
                while (i + 1 < L && xm[i + 1] <= df[k]) i++;
                inds.push_back(i);
            }
    }
    std::vector<double> tmpl(tt > 0 ? 2 * tt : 1, 0.0);
    const double ts[B] = {0.0, 0.0}, te[B] = {0.0, 0.0};
    orc_interp_tdi(tmpl.data(), toff.data(), df.data(), freqs.data(), buf.data(), sc1.data(), sc2.data(),
                   sc3.data(), ts, te, L, B, M, inds.data(), ioff.data(), st.data(), len.data());
    fails += check_finite("template build", tmpl);

    // inner products: direct on the dense band, heterodyned on the sparse grid
    std::vector<double> data(6L * N), w(3L * N);
    for (long k = 0; k < 6L * N; k++) data[k] = std::sin(0.37 * k);
    for (long k = 0; k < 3L * N; k++) w[k] = 1.0 + 0.1 * std::cos(0.11 * k);
    std::vector<double> dh(2 * B, 0.0), hh(2 * B, 0.0);
    orc_direct_like(dh.data(), hh.data(), data.data(), w.data(), tmpl.data(), toff.data(), st.data(), len.data(), N,
                    B);
    fails += check_finite("direct_like", dh) + check_finite("direct_like", hh);

    std::vector<double> sparse(2L * B * 3 * L, 0.0);
    orc_direct_sum(sparse.data(), buf.data(), B, L, 3, M, ts, te);
    std::vector<double> res(2L * L * 3 * B), dc(2L * 4 * 3 * L), nodes(L);
    for (long k = 0; k < (long)res.size(); k++) res[k] = 1.0 + 1e-3 * std::sin(0.01 * k);
    for (long k = 0; k < (long)dc.size(); k++) dc[k] = std::cos(0.02 * k);
    for (int i = 0; i < L; i++) nodes[i] = 1e-4 * (1.0 + i);
    std::vector<double> o1(2 * B), o2(2 * B);
    orc_hdyn(o1.data(), o2.data(), res.data(), dc.data(), nodes.data(), B, L, 3, 1);
    fails += check_finite("direct_sum", sparse) + check_finite("hdyn", o1) + check_finite("hdyn", o2);

    std::printf(fails ? "sanitize driver: %d failures\n" : "sanitize driver: OK\n", fails);
    return fails ? 1 : 0;
}
