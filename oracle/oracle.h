// FP64 CPU oracle for the PhenomHM -> LISA TDI response -> spline ->
// likelihood path.
//
// TEST INFRASTRUCTURE ONLY: this library is the checker that the HIP kernels
// are compared against (and the CPU baseline that bench.py times). Nothing in
// the shipped package links it.
//
// Every entry point restates a reference function in plain C++ and names the
// reference file:line it follows (paths relative to src/bbhx/ of the
// reference). All arrays use the reference's flat layouts:
//   waveform/response buffer  ((p*B + b)*M + m)*L + i     p = 0..8
//   per-mode frequencies      (b*M + m)*L + i
//   complex arrays            interleaved (re, im) doubles
#pragma once

#include <complex>

namespace orc {

typedef std::complex<double> cplx;

// Reference constants (cutils/constants.h, utils/constants.py).
constexpr double MSUN_SI = 1.98848e+30;
constexpr double YRSID_SI = 31558149.763545603;
constexpr double AU_SI = 149597870700.0;
constexpr double C_SI = 299792458.;
constexpr double MTSUN_SI = 4.925491025873693e-06;
constexpr double MRSUN_SI = 1476.6250615036158;
constexpr double PC_SI = 3.0856775814913674e+16;
constexpr double PI = 3.141592653589793238462643383279502884;
constexpr double PI_2 = 1.570796326794896619231321691639751442;
constexpr double PI_4 = 0.785398163397448309615660845819875721;
constexpr double GAMMA_E = 0.577215664901532860606512090082402431;
constexpr double SQRT2 = 1.414213562373095048801688724209698079;
constexpr double INVSQRT3 = 0.577350269189625764509148780501957455;
constexpr double INVSQRT6 = 0.408248290463863016366214012450981898;
constexpr double Omega0 = 1.9909865927683788e-07;

}  // namespace orc

extern "C" {

// ---- PhenomHM / PhenomD (cutils/PhenomHMWaveform.cu) ----
void orc_waveform_amp_phase(double* out, const int* ells, const int* mms, const double* freqs,
                            const double* m1_SI, const double* m2_SI, const double* chi1z,
                            const double* chi2z, const double* distance, const double* f_ref,
                            int numModes, int length, int numBinAll, const double* Mf_RD_all,
                            const double* Mf_DM_all, int run_phenomd);
void orc_get_phenomhm_ringdown_frequencies(double* fringdown, double* fdamp, const double* m1,
                                           const double* m2, const double* chi1z,
                                           const double* chi2z, const int* ells, const int* mms,
                                           int numModes, int numBinAll);
void orc_get_phenomd_ringdown_frequencies(double* fringdown, double* fdamp, const double* m1,
                                          const double* m2, const double* chi1z,
                                          const double* chi2z, int numBinAll, const double* y_rd,
                                          const double* c1_rd, const double* c2_rd,
                                          const double* c3_rd, const double* y_dm,
                                          const double* c1_dm, const double* c2_dm,
                                          const double* c3_dm, double dspin);

// ---- LISA fast-FD TDI response (cutils/Response.cu) with tabulated orbits ----
// Orbit tables follow the lisaanalysistools Orbits layout
// (response/fastfdresponse.py:107-134): x_arr (N, 3, 3) spacecraft positions,
// n_arr (N, 6, 3) link unit vectors, links/sc_r/sc_e (6,), uniform time step dt.
void orc_lisa_response(double* buffer, const int* ells, const int* mms, const double* freqs,
                       const double* phi_ref, const double* inc, const double* lam,
                       const double* beta, const double* psi, int TDItag, int rescaled,
                       int tdi2, int numModes, int length, int numBinAll, int includesAmps,
                       double orbit_dt, int orbit_N, const double* n_arr, const double* x_arr,
                       const int* links, double armlength);

// ---- spline (cutils/Interpolate.cu), solved with a pivoting band solver ----
void orc_interpolate(const double* x, const double* y, double* B, double* c1, double* c2,
                     double* c3, int length, int numInterpParams, int numModes, int numBinAll);

// ---- template build and inner products (cutils/WaveformBuild.cu, Likelihood.cu) ----
void orc_direct_sum(double* templateChannels, const double* buffer, int numBinAll,
                    int data_length, int nChannels, int numModes, const double* t_start,
                    const double* t_end);
void orc_interp_tdi(double* templates, const long* template_offsets, const double* dataFreqs,
                    const double* freqs, const double* y, const double* c1, const double* c2,
                    const double* c3, const double* t_start, const double* t_end, int length,
                    int numBinAll, int numModes, const int* inds, const long* inds_offsets,
                    const int* inds_start, const int* ind_lengths);
void orc_hdyn(double* likeOut1, double* likeOut2, const double* templateChannels,
              const double* dataConstants, const double* dataFreqs, int numBinAll,
              int data_length, int nChannels, int include_last_segment);
void orc_direct_like(double* d_h, double* h_h, const double* dataChannels,
                     const double* noise_weight_times_df, double* templates,
                     const long* template_offsets, const int* inds_start,
                     const int* ind_lengths, int data_stream_length, int numBinAll);

}  // extern "C"
