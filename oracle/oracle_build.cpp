// FP64 CPU restatement of the template build (direct mode sum, dense spline
// evaluation) and of the direct and heterodyned inner products.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Follows cutils/WaveformBuild.cu and cutils/Likelihood.cu of the reference.

#include "oracle.h"

#include <cmath>

using orc::cplx;

namespace {

// combine_information, WaveformBuild.cu:48-60: conj(T_ch * A e^{i phi}) inside
// the time window [t_start, t_end] (t_end <= 0 means open-ended), else 0.
inline void combine(cplx* ch, double amp, double phase, double tf, const cplx T[3], double t_start,
                    double t_end)
{
    if ((tf >= t_start) && ((tf <= t_end) || (t_end <= 0.0))) {
        cplx ap = amp * std::exp(cplx(0.0, phase));
        for (int c = 0; c < 3; c++) ch[c] = std::conj(T[c] * ap);
    }
}

}  // namespace

extern "C" {

// direct_sum -> fill_waveform, WaveformBuild.cu:194-296. Output (B, 3, L) complex.
void orc_direct_sum(double* templateChannels, const double* buf, int numBinAll, int data_length,
                    int /*nChannels*/, int numModes, const double* t_start, const double* t_end)
{
    cplx* out = reinterpret_cast<cplx*>(templateChannels);
    const long nper = (long)numBinAll * numModes * data_length;
#pragma omp parallel for schedule(static)
    for (int b = 0; b < numBinAll; b++) {
        for (int i = 0; i < data_length; i++) {
            cplx acc[3] = {0.0, 0.0, 0.0};
            for (int m = 0; m < numModes; m++) {
                long idx = ((long)b * numModes + m) * data_length + i;
                cplx T[3] = {cplx(buf[3 * nper + idx], buf[4 * nper + idx]),
                             cplx(buf[5 * nper + idx], buf[6 * nper + idx]),
                             cplx(buf[7 * nper + idx], buf[8 * nper + idx])};
                cplx ch[3] = {0.0, 0.0, 0.0};
                combine(ch, buf[idx], buf[nper + idx], buf[2 * nper + idx], T, t_start[b], t_end[b]);
                for (int c = 0; c < 3; c++) acc[c] += ch[c];
            }
            for (int c = 0; c < 3; c++) out[((long)b * 3 + c) * data_length + i] = acc[c];
        }
    }
}

// InterpTDI -> TDI, WaveformBuild.cu:68-189. Templates are (3, n_b) complex per
// binary at complex offset template_offsets[b]; segment indices (M, n_b) int32
// per binary at inds_offsets[b].
void orc_interp_tdi(double* templates, const long* template_offsets, const double* dataFreqs,
                    const double* freqs, const double* y, const double* c1, const double* c2,
                    const double* c3, const double* t_start, const double* t_end, int length,
                    int numBinAll, int numModes, const int* inds, const long* inds_offsets,
                    const int* inds_start, const int* ind_lengths)
{
    const long nper = (long)numBinAll * numModes * length;
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < numBinAll; b++) {
        cplx* tmpl = reinterpret_cast<cplx*>(templates) + template_offsets[b];
        const int* ind_b = inds + inds_offsets[b];
        const int nb = ind_lengths[b];
        for (int i = 0; i < nb; i++) {
            double f = dataFreqs[i + inds_start[b]];
            cplx acc[3] = {0.0, 0.0, 0.0};
            for (int m = 0; m < numModes; m++) {
                int k = ind_b[m * nb + i];
                if ((k <= -1) || (k >= length - 1)) continue;
                long base = ((long)b * numModes + m) * length;
                double x = f - freqs[base + k];
                double x2 = x * x, x3 = x * x2;
                double v[9];
                for (int p = 0; p < 9; p++) {
                    long q = p * nper + base + k;
                    v[p] = y[q] + c1[q] * x + c2[q] * x2 + c3[q] * x3;
                }
                cplx T[3] = {cplx(v[3], v[4]), cplx(v[5], v[6]), cplx(v[7], v[8])};
                cplx ch[3] = {0.0, 0.0, 0.0};
                combine(ch, v[0], v[1], v[2], T, t_start[b], t_end[b]);
                for (int c = 0; c < 3; c++) acc[c] += ch[c];
            }
            for (int c = 0; c < 3; c++) tmpl[(long)c * nb + i] = acc[c];
        }
    }
}

// hdyn -> hdynLikelihood, Likelihood.cu:21-201. include_last_segment = 0 gives
// the reference CPU loop (stops at jj < K-1, Likelihood.cu:150); 1 gives the
// reference GPU kernel, which also uses the last segment (Likelihood.cu:57-75).
void orc_hdyn(double* likeOut1, double* likeOut2, const double* templateChannels,
              const double* dataConstants, const double* dataFreqs, int numBinAll,
              int data_length, int nChannels, int include_last_segment)
{
    const cplx* r = reinterpret_cast<const cplx*>(templateChannels);
    const cplx* dc = reinterpret_cast<const cplx*>(dataConstants);
    cplx* o1 = reinterpret_cast<cplx*>(likeOut1);
    cplx* o2 = reinterpret_cast<cplx*>(likeOut2);
    const int jend = include_last_segment ? data_length : data_length - 1;
#pragma omp parallel for schedule(static)
    for (int b = 0; b < numBinAll; b++) {
        cplx l1 = 0.0, l2 = 0.0;
        for (int ch = 0; ch < nChannels; ch++) {
            cplx prev = 0.0;
            double fprev = 0.0;
            for (int j = 0; j < jend; j++) {
                double f = dataFreqs[j];
                cplx cur = r[((long)j * nChannels + ch) * numBinAll + b];
                if ((fprev != 0.0) && (j > 0)) {
                    cplx A0 = dc[(0 * nChannels + ch) * data_length + j];
                    cplx A1 = dc[(1 * nChannels + ch) * data_length + j];
                    cplx B0 = dc[(2 * nChannels + ch) * data_length + j];
                    cplx B1 = dc[(3 * nChannels + ch) * data_length + j];
                    cplx r1 = (cur - prev) / (f - fprev);
                    double fmid = (f + fprev) / 2.0;
                    cplx r0 = cur - r1 * (f - fmid);
                    cplx r1c = std::conj(r1);
                    l1 += A0 * std::conj(r0) + A1 * r1c;
                    double mag = std::hypot(r0.real(), r0.imag());
                    l2 += B0 * (mag * mag) + 2. * B1 * (r0 * r1c).real();
                }
                prev = cur;
                fprev = f;
            }
        }
        o1[b] = l1;
        o2[b] = l2;
    }
}

// direct_like, Likelihood.cu:357-375 and :478-529 (CPU branch). The templates
// are multiplied by the noise weights in place, as in the reference; d_h and
// h_h accumulate (+=) into what the caller passes.
void orc_direct_like(double* d_h, double* h_h, const double* dataChannels,
                     const double* noise_weight_times_df, double* templates,
                     const long* template_offsets, const int* inds_start,
                     const int* ind_lengths, int data_stream_length, int numBinAll)
{
    const cplx* d = reinterpret_cast<const cplx*>(dataChannels);
    cplx* dh = reinterpret_cast<cplx*>(d_h);
    cplx* hh = reinterpret_cast<cplx*>(h_h);
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < numBinAll; b++) {
        cplx* t = reinterpret_cast<cplx*>(templates) + template_offsets[b];
        const int nb = ind_lengths[b], st = inds_start[b];
        for (int i = 0; i < nb; i++)
            for (int c = 0; c < 3; c++) t[c * nb + i] = t[c * nb + i] * noise_weight_times_df[c * data_stream_length + st + i];
        for (int c = 0; c < 3; c++) {
            cplx s_dh = 0.0, s_hh = 0.0;
            for (int i = 0; i < nb; i++) s_dh += std::conj(d[c * data_stream_length + st + i]) * t[c * nb + i];
            dh[b] += 4.0 * s_dh;
            for (int i = 0; i < nb; i++) s_hh += std::conj(t[c * nb + i]) * t[c * nb + i];
            hh[b] += 4.0 * s_hh;
        }
    }
}

}  // extern "C"
