// Counted run of the oracle's heterodyned-likelihood forward path
// (TEST INFRASTRUCTURE ONLY): the oracle sources are compiled here with
// `double` replaced by the counting type of opcount.h, and opc_hdyn_forward
// runs, for a batch of templates on the sparse nodes, exactly the chain that
// HeterodynedLikelihood.get_ll runs (likelihood.py:550-648):
//   ringdown frequencies -> PhenomHM amp/phase/tf (PhenomHMWaveform.cu)
//   -> phase += 2 pi f (t_ref - t0), tf += t_ref (phenomhm.py:487-507)
//   -> TDI response (Response.cu) -> mode sum (WaveformBuild.cu:194-296)
//   -> r = h / h0 (likelihood.py:603) -> heterodyned sums (Likelihood.cu:125-185)
// and returns the operation counts (see opcount.h for the counting rules).

#include "opcount.h"

namespace opc {
Counts g;
}

#define double opc::Real
#include "oracle_build.cpp"
#include "oracle_phenomhm.cpp"
#include "oracle_response.cpp"
#undef double

extern "C" {

// Operation names in the order of opc::Op.
const char* opc_op_names(void)
{
    return "add,mul,div,sqrt,sin,cos,exp,log,pow,cbrt,atan,atan2,hypot,floor,fabs";
}

int opc_num_ops(void) { return opc::NOPS; }

// params: 11 rows x B (m1, m2 [Msun], chi1, chi2, dist [m], phi_ref, inc, lam,
// beta, psi, t_ref [s]); freqs: K sparse nodes; h0: (3, K) complex reference
// template; dc: (4, 3, K) complex heterodyne constants; qnm: 8 x table_len
// PhenomD ringdown coefficients (y_rd, c1_rd, c2_rd, c3_rd, y_dm, ...);
// t_start, t_end: (B,) observation window in the SSB frame.
// counts: (NOPS,) totals over the batch; ll: (B,) log-likelihoods (-1/2 (h_h - 2 Re d_h)
// without d_d, for checking the counted run against the uncounted oracle).
int opc_hdyn_forward(double* counts, double* ll, const double* params, int B, const double* freqs, int K,
                     const double* h0, const double* dc, const double* qnm, int table_len, double dspin,
                     const double* t_start, const double* t_end, double initial_t_val, double orbit_dt,
                     int orbit_N, const double* n_arr, const double* x_arr, const int* links,
                     double armlength, int tdi2)
{
    using opc::Real;
    const int M = 6;
    const int ells[M] = {2, 3, 4, 2, 3, 4};
    const int mms[M] = {2, 3, 4, 1, 2, 3};
    const long nper = (long)B * M * K;
    std::vector<Real> P(11 * (size_t)B), m1(B), m2(B), m1s(B), m2s(B), fref(B, Real(0.0));
    for (size_t i = 0; i < P.size(); i++) P[i] = Real(params[i]);
    std::vector<Real> fq(nper), buf(9 * nper, Real(0.0)), Q(8 * (size_t)table_len);
    std::vector<Real> frd(B * (M + 1)), fdm(B * (M + 1)), hr(B * M), hd(B * M), dr(B), dd(B);
    std::vector<Real> tmpl(2 * 3 * (size_t)B * K), resid(2 * 3 * (size_t)B * K), o1(2 * (size_t)B), o2(2 * (size_t)B);
    std::vector<Real> nA(n_arr, n_arr + (size_t)orbit_N * 18), xA(x_arr, x_arr + (size_t)orbit_N * 9);
    std::vector<Real> dcR(dc, dc + 2 * 4 * 3 * (size_t)K), fR(freqs, freqs + K), ts(t_start, t_start + B),
        te(t_end, t_end + B);
    for (size_t i = 0; i < Q.size(); i++) Q[i] = Real(qnm[i]);
    for (int b = 0; b < B; b++)
        for (int m = 0; m < M; m++)
            for (int i = 0; i < K; i++) fq[((long)b * M + m) * K + i] = Real(freqs[i]);
    std::memset(&opc::g, 0, sizeof(opc::g));

    // phenomhm.py:175-199 (m1 >= m2 swap is a no-op for the sampled parameters) and :395-463
    for (int b = 0; b < B; b++) {
        m1[b] = P[0 * B + b];
        m2[b] = P[1 * B + b];
    }
    const int L = table_len;
    orc_get_phenomd_ringdown_frequencies(dr.data(), dd.data(), m1.data(), m2.data(), &P[2 * B], &P[3 * B], B,
                                         &Q[0 * L], &Q[1 * L], &Q[2 * L], &Q[3 * L], &Q[4 * L], &Q[5 * L],
                                         &Q[6 * L], &Q[7 * L], Real(dspin));
    orc_get_phenomhm_ringdown_frequencies(hr.data(), hd.data(), m1.data(), m2.data(), &P[2 * B], &P[3 * B], ells,
                                          mms, M, B);
    for (int b = 0; b < B; b++) {
        for (int m = 0; m < M; m++) {
            frd[b * (M + 1) + m] = hr[b * M + m];
            fdm[b * (M + 1) + m] = hd[b * M + m];
        }
        frd[b * (M + 1) + M] = dr[b];
        fdm[b * (M + 1) + M] = dd[b];
        m1s[b] = m1[b] * orc::MSUN_SI;
        m2s[b] = m2[b] * orc::MSUN_SI;
    }
    orc_waveform_amp_phase(buf.data(), ells, mms, fq.data(), m1s.data(), m2s.data(), &P[2 * B], &P[3 * B], &P[4 * B],
                           fref.data(), M, K, B, frd.data(), fdm.data(), 0);
    for (int b = 0; b < B; b++) {
        const Real tref = P[10 * B + b];
        for (long q = (long)b * M * K; q < (long)(b + 1) * M * K; q++) {
            buf[nper + q] += fq[q] * (tref - Real(initial_t_val)) * Real(2.0) * Real(orc::PI);
            buf[2 * nper + q] += tref;
        }
    }
    orc_lisa_response(buf.data(), ells, mms, fq.data(), &P[5 * B], &P[6 * B], &P[7 * B], &P[8 * B], &P[9 * B], 2, 0,
                      tdi2, M, K, B, 1, Real(orbit_dt), orbit_N, nA.data(), xA.data(), links, Real(armlength));
    orc_direct_sum(tmpl.data(), buf.data(), B, K, 3, M, ts.data(), te.data());
    // r = h / h0 into the hdyn layout (j * 3 + ch) * B + b
    typedef std::complex<Real> cr;
    const cr* tc = reinterpret_cast<const cr*>(tmpl.data());
    cr* rc = reinterpret_cast<cr*>(resid.data());
    for (int b = 0; b < B; b++)
        for (int c = 0; c < 3; c++)
            for (int j = 0; j < K; j++) {
                const cr hh0(Real(h0[2 * (c * K + j)]), Real(h0[2 * (c * K + j) + 1]));
                rc[((long)j * 3 + c) * B + b] = tc[((long)b * 3 + c) * K + j] / hh0;
            }
    orc_hdyn(o1.data(), o2.data(), resid.data(), dcR.data(), fR.data(), B, K, 3, 1);
    for (int b = 0; b < B; b++) ll[b] = -0.5 * (o2[2 * b].v - 2.0 * o1[2 * b].v);
    for (int o = 0; o < opc::NOPS; o++) counts[o] = double(opc::g.n[o]);
    return 0;
}

// Counted dense evaluation and direct inner products of one batch
// (TemplateInterpFD + Likelihood.get_ll, waveformbuild.py:89-239,
// likelihood.py:140-225): orc_interp_tdi then orc_direct_like on the same
// arguments the uncounted oracle receives (see oracle/__init__.py). Returns
// the counts and d_h, h_h (B,) complex for checking against the uncounted run.
int opc_interp_direct(double* counts, double* d_h, double* h_h, const double* dataFreqs, int data_length,
                      const double* freqs, const double* y, const double* c1, const double* c2, const double* c3,
                      const double* t_start, const double* t_end, int length, int numBinAll, int numModes,
                      const int* inds, const long* inds_offsets, const int* inds_start, const int* ind_lengths,
                      const double* dataChannels, const double* noise)
{
    using opc::Real;
    const long nper = (long)numBinAll * numModes * length;
    auto R = [](const double* p, size_t n) { return std::vector<Real>(p, p + n); };
    std::vector<Real> df = R(dataFreqs, data_length), fr = R(freqs, nper), Y = R(y, 9 * nper), C1 = R(c1, 9 * nper),
                      C2 = R(c2, 9 * nper), C3 = R(c3, 9 * nper), ts = R(t_start, numBinAll),
                      te = R(t_end, numBinAll), D = R(dataChannels, 6 * (size_t)data_length),
                      W = R(noise, 3 * (size_t)data_length);
    std::vector<long> toff(numBinAll);
    long tot = 0;
    for (int b = 0; b < numBinAll; b++) {
        toff[b] = tot;
        tot += 3L * ind_lengths[b];
    }
    std::vector<Real> tmpl(2 * (size_t)tot, Real(0.0)), dh(2 * (size_t)numBinAll, Real(0.0)),
        hh(2 * (size_t)numBinAll, Real(0.0));
    std::memset(&opc::g, 0, sizeof(opc::g));
    orc_interp_tdi(tmpl.data(), toff.data(), df.data(), fr.data(), Y.data(), C1.data(), C2.data(), C3.data(),
                   ts.data(), te.data(), length, numBinAll, numModes, inds, inds_offsets, inds_start, ind_lengths);
    orc_direct_like(dh.data(), hh.data(), D.data(), W.data(), tmpl.data(), toff.data(), inds_start, ind_lengths,
                    data_length, numBinAll);
    for (int b = 0; b < 2 * numBinAll; b++) {
        d_h[b] = dh[b].v;
        h_h[b] = hh[b].v;
    }
    for (int o = 0; o < opc::NOPS; o++) counts[o] = double(opc::g.n[o]);
    return 0;
}

}  // extern "C"
