// PhenomHM / PhenomD kernels and their C-ABI entry points.
//
//  k_phenomd_ringdown   one lane per binary: PhenomD fRD, fDM from the QNM table
//  k_phenomhm_ringdown  one lane per binary: per-mode QNM fRD, fDM (CW07102016 fits)
//  k_phm_setup          one lane per (binary, mode): coefficient record PhmRec
//  k_phm_eval           one lane per (binary, mode, frequency): amp, phase, t(f)
//
// Layouts are the reference's (PhenomHMWaveform.cu:2891-3083): output buffer
// slabs amp | phase | tf of (B, M, L); frequencies (B, M, L); ringdown
// frequencies (B, M+1) with the PhenomD value in the last column.

#include "bbhx_hip.h"
#include "common.h"
#include "phenomhm_device.h"
#include "workspace.h"

using namespace bbhx;
using namespace bbhx::phm;

namespace {

constexpr int EVAL_THREADS = 256;

__global__ __launch_bounds__(256) void k_phenomd_ringdown(double* __restrict__ fringdown, double* __restrict__ fdamp,
                                   const double* __restrict__ m1, const double* __restrict__ m2,
                                   const double* __restrict__ chi1z, const double* __restrict__ chi2z,
                                   int numBinAll, const double* __restrict__ y_rd, const double* __restrict__ c1_rd,
                                   const double* __restrict__ c2_rd, const double* __restrict__ c3_rd,
                                   const double* __restrict__ y_dm, const double* __restrict__ c1_dm,
                                   const double* __restrict__ c2_dm, const double* __restrict__ c3_dm, double dspin,
                                   int table_len)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= numBinAll) return;
    double finmass, finspin;
    final_state(m1[b], m2[b], chi1z[b], chi2z[b], finmass, finspin);
    const double lowest = -1.0;
    int k = int((finspin - lowest) / dspin);
    // the reference indexes without a bound; keep the lookup inside the table
    k = k < 0 ? 0 : (k > table_len - 1 ? table_len - 1 : k);
    double x = finspin - (dspin * k + lowest);
    double x2 = x * x, x3 = x2 * x;
    fringdown[b] = (y_rd[k] + c1_rd[k] * x + c2_rd[k] * x2 + c3_rd[k] * x3) / finmass;
    fdamp[b] = (y_dm[k] + c1_dm[k] * x + c2_dm[k] * x2 + c3_dm[k] * x3) / finmass;
}

__global__ void k_phenomhm_ringdown(double* __restrict__ fringdown, double* __restrict__ fdamp,
                                    const double* __restrict__ m1, const double* __restrict__ m2,
                                    const double* __restrict__ chi1z, const double* __restrict__ chi2z,
                                    const int* __restrict__ ells, const int* __restrict__ mms, int numModes,
                                    int numBinAll)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= numBinAll) return;
    double finmass, finspin;
    final_state(m1[b], m2[b], chi1z[b], chi2z[b], finmass, finspin);
    for (int mi = 0; mi < numModes; mi++) {
        double fr, fd;
        hm_ringdown(ells[mi], mms[mi], finmass, finspin, fr, fd);
        fringdown[b * numModes + mi] = fr;
        fdamp[b * numModes + mi] = fd;
    }
}

__global__ __launch_bounds__(64) void k_phm_setup(PhmRec* __restrict__ recs, const int* __restrict__ ells, const int* __restrict__ mms,
                            const double* __restrict__ m1_SI, const double* __restrict__ m2_SI,
                            const double* __restrict__ chi1z, const double* __restrict__ chi2z,
                            const double* __restrict__ distance, int numModes, int numBinAll,
                            const double* __restrict__ Mf_RD_all, const double* __restrict__ Mf_DM_all,
                            int run_phenomd)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= numBinAll * numModes) return;
    int b = idx / numModes, mi = idx % numModes;
    const int ncol = run_phenomd ? numModes : numModes + 1;
    const double* rd = Mf_RD_all + (long)b * ncol;
    const double* dm = Mf_DM_all + (long)b * ncol;
    const double rdD = run_phenomd ? rd[0] : rd[numModes];
    const double dmD = run_phenomd ? dm[0] : dm[numModes];
    phm_setup(recs[idx], m1_SI[b], m2_SI[b], chi1z[b], chi2z[b], distance[b], ells[mi], mms[mi], rd[mi], dm[mi], rdD,
              dmD, run_phenomd != 0);
}

__global__ __launch_bounds__(EVAL_THREADS) void k_phm_eval(double* __restrict__ out, const PhmRec* __restrict__ recs,
                                                           const double* __restrict__ freqs, int numModes,
                                                           int length, long nper)
{
    const int i = blockIdx.x * EVAL_THREADS + threadIdx.x;
    const int mi = blockIdx.y, b = blockIdx.z;
    if (i >= length) return;
    const PhmRec& r = recs[b * numModes + mi];
    const long idx = ((long)b * numModes + mi) * length + i;
    double amp, phase, tf;
    phm_eval(r, freqs[idx], amp, phase, tf);
    out[idx] = amp;
    out[nper + idx] = phase;
    out[2 * nper + idx] = tf;
}

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int launch_status()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? OK : ERR_LAUNCH;
}

}  // namespace

extern "C" {

int bbhx_hip_get_phenomd_ringdown_frequencies(double* fringdown, double* fdamp, const double* m1, const double* m2,
                                              const double* chi1z, const double* chi2z, int numBinAll,
                                              const double* y_rd, const double* c1_rd, const double* c2_rd,
                                              const double* c3_rd, const double* y_dm, const double* c1_dm,
                                              const double* c2_dm, const double* c3_dm, double dspin,
                                              int table_len, void* stream)
{
    if (numBinAll < 0 || table_len < 2 || !(dspin > 0.0)) return ERR_ARGS;
    if (numBinAll == 0) return OK;
    if (!fringdown || !fdamp || !m1 || !m2 || !chi1z || !chi2z || !y_rd || !c1_rd || !c2_rd || !c3_rd || !y_dm ||
        !c1_dm || !c2_dm || !c3_dm)
        return ERR_ARGS;
    hipLaunchKernelGGL(k_phenomd_ringdown, dim3(ceil_div(numBinAll, 256)), dim3(256), 0, S(stream), fringdown, fdamp,
                       m1, m2, chi1z, chi2z, numBinAll, y_rd, c1_rd, c2_rd, c3_rd, y_dm, c1_dm, c2_dm, c3_dm, dspin,
                       table_len);
    return launch_status();
}

int bbhx_hip_get_phenomhm_ringdown_frequencies(double* fringdown, double* fdamp, const double* m1, const double* m2,
                                               const double* chi1z, const double* chi2z, const int* ells,
                                               const int* mms, int numModes, int numBinAll, void* stream)
{
    if (numBinAll < 0 || numModes < 1 || numModes > MAX_MODES) return ERR_ARGS;
    if (numBinAll == 0) return OK;
    if (!fringdown || !fdamp || !m1 || !m2 || !chi1z || !chi2z || !ells || !mms) return ERR_ARGS;
    hipLaunchKernelGGL(k_phenomhm_ringdown, dim3(ceil_div(numBinAll, 256)), dim3(256), 0, S(stream), fringdown,
                       fdamp, m1, m2, chi1z, chi2z, ells, mms, numModes, numBinAll);
    return launch_status();
}

int bbhx_hip_waveform_amp_phase(double* waveformOut, const int* ells, const int* mms, const double* freqs,
                                const double* m1_SI, const double* m2_SI, const double* chi1z, const double* chi2z,
                                const double* distance, const double* f_ref, int numModes, int length,
                                int numBinAll, const double* Mf_RD_lm_all, const double* Mf_DM_lm_all,
                                int run_phenomd, void* stream)
{
    (void)f_ref;  // replaced by the (2,2) amplitude peak, as in the reference
    if (numBinAll < 0 || length < 1 || numModes < 1 || numModes > MAX_MODES) return ERR_ARGS;
    if (run_phenomd && numModes != 1) return ERR_ARGS;
    if (numBinAll == 0) return OK;
    if (!waveformOut || !ells || !mms || !freqs || !m1_SI || !m2_SI || !chi1z || !chi2z || !distance ||
        !Mf_RD_lm_all || !Mf_DM_lm_all)
        return ERR_ARGS;
    if ((long)numBinAll * numModes > 65535L * 65535L) return ERR_ARGS;
    const int nrec = numBinAll * numModes;
    PhmRec* recs = static_cast<PhmRec*>(workspace_get(WS_PHM_REC, sizeof(PhmRec) * (size_t)nrec, S(stream)));
    if (!recs) return ERR_ALLOC;
    hipLaunchKernelGGL(k_phm_setup, dim3(ceil_div(nrec, 64)), dim3(64), 0, S(stream), recs, ells, mms, m1_SI, m2_SI,
                       chi1z, chi2z, distance, numModes, numBinAll, Mf_RD_lm_all, Mf_DM_lm_all, run_phenomd);
    int st = launch_status();
    if (st != OK) return st;
    if (numBinAll > 65535) {
        // grid.z is limited; split the binaries into chunks
        for (int b0 = 0; b0 < numBinAll; b0 += 65535) {
            int nb = numBinAll - b0 < 65535 ? numBinAll - b0 : 65535;
            long off = (long)b0 * numModes * length;
            hipLaunchKernelGGL(k_phm_eval, dim3(ceil_div(length, EVAL_THREADS), numModes, nb), dim3(EVAL_THREADS), 0,
                               S(stream), waveformOut + off, recs + (long)b0 * numModes, freqs + off, numModes, length,
                               (long)numBinAll * numModes * length);
        }
    } else {
        hipLaunchKernelGGL(k_phm_eval, dim3(ceil_div(length, EVAL_THREADS), numModes, numBinAll), dim3(EVAL_THREADS),
                           0, S(stream), waveformOut, recs, freqs, numModes, length,
                           (long)numBinAll * numModes * length);
    }
    return launch_status();
}

}  // extern "C"
