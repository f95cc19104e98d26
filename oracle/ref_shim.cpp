// C-ABI export layer over the reference's own CPU code paths.
//
// TEST INFRASTRUCTURE ONLY. This file is compiled together with the
// reference sources cutils/PhenomHMWaveform.cu, cutils/WaveformBuild.cu and
// cutils/Likelihood.cu (as C++, i.e. their non-CUDA branches) by
// oracle/Makefile into oracle/_ref/libbbhx_ref.so. It contains no reference
// code: it only forwards plain pointers to the reference entry points declared
// in the reference headers, so that ctypes can call them without C++ name
// mangling. cmplx arrays are passed as interleaved (re, im) doubles.
//
// Not built: cutils/Interpolate.cu (needs lapacke.h, absent) and
// cutils/Response.cu (needs lisaanalysistools' Detector.hpp, absent).

#include "PhenomHMWaveform.hh"
#include "WaveformBuild.hh"
#include "Likelihood.hh"

extern "C" {

// PhenomHMWaveform.cu:2998
void ref_waveform_amp_phase(double* out, int* ells, int* mms, double* freqs,
                            double* m1_SI, double* m2_SI, double* chi1z, double* chi2z,
                            double* distance, double* f_ref, int numModes, int length,
                            int numBinAll, double* Mf_RD, double* Mf_DM, int run_phenomd)
{
    waveform_amp_phase(out, ells, mms, freqs, m1_SI, m2_SI, chi1z, chi2z, distance, f_ref,
                       numModes, length, numBinAll, Mf_RD, Mf_DM, run_phenomd);
}

// PhenomHMWaveform.cu:1879
void ref_get_phenomhm_ringdown_frequencies(double* fringdown, double* fdamp, double* m1,
                                           double* m2, double* chi1z, double* chi2z,
                                           int* ells, int* mms, int numModes, int numBinAll)
{
    get_phenomhm_ringdown_frequencies_wrap(fringdown, fdamp, m1, m2, chi1z, chi2z, ells, mms,
                                           numModes, numBinAll);
}

// PhenomHMWaveform.cu:1995
void ref_get_phenomd_ringdown_frequencies(double* fringdown, double* fdamp, double* m1,
                                          double* m2, double* chi1z, double* chi2z,
                                          int numBinAll, double* y_rd, double* c1_rd,
                                          double* c2_rd, double* c3_rd, double* y_dm,
                                          double* c1_dm, double* c2_dm, double* c3_dm,
                                          double dspin)
{
    get_phenomd_ringdown_frequencies_wrap(fringdown, fdamp, m1, m2, chi1z, chi2z, numBinAll,
                                          y_rd, c1_rd, c2_rd, c3_rd, y_dm, c1_dm, c2_dm,
                                          c3_dm, dspin);
}

// WaveformBuild.cu:281
void ref_direct_sum(double* templateChannels, double* bbh_buffer, int numBinAll,
                    int data_length, int nChannels, int numModes, double* t_start,
                    double* t_end)
{
    direct_sum(reinterpret_cast<cmplx*>(templateChannels), bbh_buffer, numBinAll, data_length,
               nChannels, numModes, t_start, t_end);
}

// WaveformBuild.cu:147
void ref_InterpTDI(long* templateChannels_ptrs, double* dataFreqs, double* freqs,
                   double* propArrays, double* c1, double* c2, double* c3, double* t_start,
                   double* t_end, int length, int data_length, int numBinAll, int numModes,
                   long* inds_ptrs, int* inds_start, int* ind_lengths)
{
    InterpTDI(templateChannels_ptrs, dataFreqs, freqs, propArrays, c1, c2, c3, t_start, t_end,
              length, data_length, numBinAll, numModes, inds_ptrs, inds_start, ind_lengths);
}

// Likelihood.cu:187 (CPU branch: Likelihood.cu:125)
void ref_hdyn(double* likeOut1, double* likeOut2, double* templateChannels,
              double* dataConstants, double* dataFreqs, int numBinAll, int data_length,
              int nChannels)
{
    hdyn(reinterpret_cast<cmplx*>(likeOut1), reinterpret_cast<cmplx*>(likeOut2),
         reinterpret_cast<cmplx*>(templateChannels), reinterpret_cast<cmplx*>(dataConstants),
         dataFreqs, numBinAll, data_length, nChannels);
}

// Likelihood.cu:492 (CPU branch)
void ref_direct_like(double* d_h, double* h_h, double* dataChannels,
                     double* noise_weight_times_df, long* templateChannels_ptrs,
                     int* inds_start, int* ind_lengths, int data_stream_length, int numBinAll)
{
    direct_like(reinterpret_cast<cmplx*>(d_h), reinterpret_cast<cmplx*>(h_h),
                reinterpret_cast<cmplx*>(dataChannels), noise_weight_times_df,
                templateChannels_ptrs, inds_start, ind_lengths, data_stream_length, numBinAll);
}

}  // extern "C"
