/*
 * C ABI of the gfx950 (MI355X) backend for the BBHx template-to-likelihood path.
 *
 * Every entry point:
 *   - takes device pointers (unless stated "host") and plain sizes;
 *   - is asynchronous on 'stream' (a hipStream_t passed as void*; NULL = the
 *     default stream) and returns as soon as the work is queued, except where a
 *     host-side table must be uploaded first (noted below);
 *   - returns 0 on success or a nonzero status: 1 invalid arguments,
 *     2 launch/runtime error, 3 workspace allocation failure, 4 unsupported mode,
 *     5 batch above the capacity of an allocation-free entry point.
 *     It never exits the process (the reference calls exit() on errors,
 *     cutils/global.h:14-22).
 *
 * Array layouts are the reference's flat layouts:
 *   waveform / response buffer   ((p*B + b)*M + m)*L + i,  p = amp, phase, tf,
 *                                T1re, T1im, T2re, T2im, T3re, T3im
 *   frequencies                  (b*M + m)*L + i
 *   ringdown frequencies         b*(M+1) + m, PhenomD value in column M
 *   complex arrays               interleaved (re, im) doubles
 *
 * The library keeps grow-only scratch buffers (coefficient records, band
 * tables, partial sums), one set per (device, stream): calls queued on
 * different streams never share scratch memory. A buffer that is replaced by a
 * larger one is kept until bbhx_hip_release_workspace, so a launch sequence
 * captured in a graph stays valid after later, larger calls on the same
 * stream; replays of one graph must not run concurrently with each other.
 * Scratch is never allocated while the stream is being captured (that would
 * invalidate the capture; the call returns 3 instead): before capturing a
 * call, run it once eagerly on the capture stream with a batch at least as
 * large.
 *
 * Reference file:line citations are relative to src/bbhx/ of mikekatz04/BBHx.
 */
#ifndef BBHX_HIP_H
#define BBHX_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- waveform: PhenomHM / PhenomD ------------------------------------------ */

/* Replaces waveform_amp_phase_wrap (cutils/phenomhm.pyx:61) ->
 * waveform_amp_phase (cutils/PhenomHMWaveform.cu:2998).
 * Fills the amp | phase | tf slabs of waveformOut (B, M, L each). ells/mms (M,)
 * int32; masses in kg; distance in m; f_ref is ignored (the reference replaces
 * it by the (2,2) amplitude peak, PhenomHMWaveform.cu:2767). Mf_RD/DM_lm_all:
 * (B, M+1) from the two ringdown entry points below, or (B, 1) when
 * run_phenomd (then numModes must be 1). */
int bbhx_hip_waveform_amp_phase(double* waveformOut, const int* ells, const int* mms, const double* freqs,
                                const double* m1_SI, const double* m2_SI, const double* chi1z, const double* chi2z,
                                const double* distance, const double* f_ref, int numModes, int length,
                                int numBinAll, const double* Mf_RD_lm_all, const double* Mf_DM_lm_all,
                                int run_phenomd, void* stream);

/* Replaces get_phenomhm_ringdown_frequencies (cutils/phenomhm.pyx:117 ->
 * PhenomHMWaveform.cu:1817-1928). Masses in solar masses. Output (B, M). */
int bbhx_hip_get_phenomhm_ringdown_frequencies(double* fringdown, double* fdamp, const double* m1, const double* m2,
                                               const double* chi1z, const double* chi2z, const int* ells,
                                               const int* mms, int numModes, int numBinAll, void* stream);

/* Replaces get_phenomd_ringdown_frequencies (cutils/phenomhm.pyx:155 ->
 * PhenomHMWaveform.cu:1933-2062). Spline coefficient arrays of the QNM table
 * have table_len entries each (1001 for the reference's table; the index is
 * clamped into the table, the reference does not check it). Output (B,). */
int bbhx_hip_get_phenomd_ringdown_frequencies(double* fringdown, double* fdamp, const double* m1, const double* m2,
                                              const double* chi1z, const double* chi2z, int numBinAll,
                                              const double* y_rd, const double* c1_rd, const double* c2_rd,
                                              const double* c3_rd, const double* y_dm, const double* c1_dm,
                                              const double* c2_dm, const double* c3_dm, double dspin,
                                              int table_len, void* stream);

/* ---- LISA TDI response (stateful, like pyFastLISAResponse) ------------------ */

typedef struct bbhx_response bbhx_response;

/* Replaces pyFastLISAResponse.__cinit__ / __dealloc__ (cutils/lisaresponse.pyx:34-40, :117-121). */
int bbhx_hip_response_create(bbhx_response** out);
int bbhx_hip_response_destroy(bbhx_response* r);

/* Replaces pyFastLISAResponse.add_orbit_information (cutils/lisaresponse.pyx:41-71):
 * orbit tables in the lisaanalysistools Orbits layout. n_arr (N, 6, 3),
 * L_arr (N, 6) and x_arr (N, 3, 3) are device arrays the caller keeps alive;
 * links, sc_r, sc_e are 6 host ints. Positions/unit vectors are linearly
 * interpolated in time with step dt (response/fastfdresponse.py:174). */
int bbhx_hip_response_add_orbits(bbhx_response* r, double dt, int N, const double* n_arr, const double* L_arr,
                                 const double* x_arr, const int* links, const int* sc_r, const int* sc_e,
                                 double armlength);

/* Replaces pyFastLISAResponse.LISA_response_wrap (cutils/lisaresponse.pyx:78) ->
 * FastLISAResponse::LISA_response (cutils/Response.cu:671-746). TDItag 1 = XYZ,
 * 2 = AET. Writes the 6 transfer-function slabs and adds the orbital Doppler
 * phase to the phase slab in place. order_fresnel_stencil must be 0. */
int bbhx_hip_response_compute(bbhx_response* r, double* response_out, const int* ells, const int* mms,
                              const double* freqs, const double* phi_ref, const double* inc, const double* lam,
                              const double* beta, const double* psi, int TDItag, int rescaled, int tdi2,
                              int order_fresnel_stencil, int numModes, int length, int numBinAll, int includesAmps,
                              void* stream);

/* ---- spline and template build --------------------------------------------- */

/* Replaces interpolate_wrap (cutils/interp.pyx:14) -> interpolate
 * (cutils/Interpolate.cu:320-359). Not-a-knot cubic splines of the (P, B, M, L)
 * values over the (B, M, L) abscissae. B receives the knot slopes; upper_diag,
 * diag, lower_diag receive c1, c2, c3 (as in the reference); the entry at the
 * last knot of each is 0. length >= 3. */
int bbhx_hip_interpolate(const double* freqs, const double* propArrays, double* B, double* upper_diag,
                         double* diag, double* lower_diag, int length, int numInterpParams, int numModes,
                         int numBinAll, void* stream);

/* Replaces InterpTDI_wrap (cutils/bbhwaveformbuild.pyx:18) -> InterpTDI
 * (cutils/WaveformBuild.cu:147-189). templateChannels_ptrs, inds_ptrs,
 * inds_start, ind_lengths, t_start, t_end are HOST arrays of B entries (device
 * addresses of each binary's (3, n_b) complex template and (M, n_b) int32
 * segment indices). One launch for all binaries; waits for the upload of the
 * B-entry band table before returning. */
int bbhx_hip_InterpTDI(const int64_t* templateChannels_ptrs, const double* dataFreqs, const double* freqs,
                       const double* propArrays, const double* c1, const double* c2, const double* c3,
                       const double* t_start, const double* t_end, int length, int data_length, int numBinAll,
                       int numModes, const int64_t* inds_ptrs, const int* inds_start, const int* ind_lengths,
                       void* stream);

/* Replaces direct_sum_wrap (cutils/bbhwaveformbuild.pyx:37) -> direct_sum
 * (cutils/WaveformBuild.cu:194-296). Output (B, 3, L) complex; t_start/t_end (B,) device. */
int bbhx_hip_direct_sum(double* templateChannels, const double* bbh_buffer, int numBinAll, int data_length,
                        int nChannels, int numModes, const double* t_start, const double* t_end, void* stream);

/* ---- inner products --------------------------------------------------------- */

/* Replaces hdyn_wrap (cutils/bbhlikelihood.pyx:21) -> hdyn (cutils/Likelihood.cu:21-201).
 * templateChannels: residuals r = h/h0, (K, 3, B) complex; dataConstants
 * (4, 3, K) complex (A0, A1, B0, B1); outputs d_h, h_h (B,) complex.
 * include_last_segment = 1 follows the reference GPU kernel (all K-1
 * segments), 0 the reference CPU loop (stops one segment early, Likelihood.cu:150). */
int bbhx_hip_hdyn(double* likeOut1, double* likeOut2, const double* templateChannels, const double* dataConstants,
                  const double* dataFreqs, int numBinAll, int data_length, int nChannels, int include_last_segment,
                  void* stream);

/* Replaces direct_like_wrap (cutils/bbhlikelihood.pyx:42) -> direct_like
 * (cutils/Likelihood.cu:357-471). d_h, h_h: (B,) complex DEVICE arrays,
 * accumulated with += (the reference accumulates into host arrays).
 * templateChannels_ptrs, inds_start, ind_lengths: HOST arrays of B entries.
 * The templates are multiplied by the noise weights in place, as in the
 * reference. One launch for all binaries; deterministic reduction. */
int bbhx_hip_direct_like(double* d_h, double* h_h, const double* dataChannels, const double* noise_weight_times_df,
                         const int64_t* templateChannels_ptrs, const int* inds_start, const int* ind_lengths,
                         int data_stream_length, int numBinAll, void* stream);

/* ---- fused heterodyned likelihood (no counterpart in the reference) --------- */

/* One pass of HeterodynedLikelihood.get_ll (likelihood.py:550-648) without
 * materialising the waveform buffer: for each template, PhenomHM amplitude /
 * phase / t(f) (+ the t_ref shift, phenomhm.py:487-507), the TDI response, the
 * windowed mode sum (WaveformBuild.cu:48-60), r = h / h0 and the heterodyned
 * sums (Likelihood.cu:52-111), all at the K sparse nodes.
 *
 * params: (11, B) device, rows m1 [Msun], m2 [Msun], chi1z, chi2z,
 *   distance [m], phi_ref, inc, lam, beta, psi, t_ref [s] (m1 >= m2 already enforced
 *   by the host, as phenomhm.py:175-199 does).
 * t_start, t_end: (B,) device, SSB-frame window.
 * freqs (K,), h0 (3, K) complex, dataConstants (4, 3, K) complex, all device.
 * qnm: the PhenomD QNM table spline, 8 arrays of table_len entries packed
 *   contiguously (y_rd, c1_rd, c2_rd, c3_rd, y_dm, c1_dm, c2_dm, c3_dm).
 * precision: must be 0 (FP64 throughout; no reduced-precision form of the
 *   heterodyned sums is built: other values return 1).
 * Outputs d_h, h_h: (B,) complex; logL: (B,) device or null: when given, the
 *   eval kernel also writes logL[b] = -1/2 (d_d + Re h_h - 2 Re d_h)
 *   (likelihood.py:640-648, d_d = <d|d> of the data), so that a caller needs
 *   no further kernel for it. */
int bbhx_hip_hdyn_fused(double* d_h, double* h_h, const double* params, const double* t_start,
                        const double* t_end, int numBinAll, const int* ells_host, const int* mms_host,
                        int numModes, const double* freqs, const double* h0, const double* dataConstants,
                        int K, const double* qnm, int table_len, double dspin, double initial_t_val,
                        const bbhx_response* orbits, int TDItag, int rescaled, int tdi2,
                        int include_last_segment, int precision, double* logL, double d_d, void* stream);

/* No counterpart in the reference. The same computation as
 * bbhx_hip_hdyn_fused, with the same arguments and results (bit for bit), for
 * callers that cannot allocate device memory: code under HIP graph capture
 * that must not call hipMalloc, tight serving loops, harnesses that forbid
 * allocation. It calls no hipMalloc / hipFree and no allocating library
 * routine: the per-(template, mode) records and the per-node weights live in
 * a fixed array of the library, BBHX_HIP_NOALLOC_MAX_RECORDS records and
 * weights for up to 1024 nodes (about 22.6 MB) per device,
 * reserved when the library loads on a device and held until the process
 * ends. numBinAll * numModes above that returns status 5. Every call on one
 * device shares the array, which keeps the last call's records: calls that
 * overlap on one GPU (on different streams) race and must not be issued.
 * No event timing (bbhx_hip_timing_* does not see these calls). */
#define BBHX_HIP_NOALLOC_MAX_RECORDS 24576
int bbhx_hip_hdyn_fused_noalloc(double* d_h, double* h_h, const double* params, const double* t_start,
                                const double* t_end, int numBinAll, const int* ells_host, const int* mms_host,
                                int numModes, const double* freqs, const double* h0, const double* dataConstants,
                                int K, const double* qnm, int table_len, double dspin, double initial_t_val,
                                const bbhx_response* orbits, int TDItag, int rescaled, int tdi2,
                                int include_last_segment, int precision, double* logL, double d_d, void* stream);

/* Gradient of the heterodyned log-likelihood (no counterpart in the reference,
 * which has no derivatives). Same inputs as bbhx_hip_hdyn_fused (precision is
 * FP64); computes d_h, h_h as it does and grad (11, B): the derivatives of
 * logL = Re d_h - 1/2 Re h_h (- 1/2 d_d) with respect to ln(m1 + m2), q = m1/m2,
 * chi1z, chi2z, distance [m], phi_ref, cos(inc), lam, sin(beta), psi, t_ref [s]
 * (row order). The observation window t_start/t_end is held fixed. K <= 1024.
 * logL (B,) device or null, d_d: as for bbhx_hip_hdyn_fused. */
int bbhx_hip_hdyn_fused_grad(double* d_h, double* h_h, double* grad, const double* params, const double* t_start,
                             const double* t_end, int numBinAll, const int* ells_host, const int* mms_host,
                             int numModes, const double* freqs, const double* h0, const double* dataConstants, int K,
                             const double* qnm, int table_len, double dspin, double initial_t_val,
                             const bbhx_response* orbits, int TDItag, int rescaled, int tdi2,
                             int include_last_segment, double* logL, double d_d, void* stream);

/* ---- fused direct likelihood of a ragged batch (no counterpart in the reference) ---- */

/* Likelihood.get_ll (likelihood.py:140-225) with BBHWaveformFD(direct=False)
 * (waveformbuild.py:338-700) for templates whose sparse grids have different
 * lengths L_b, in one sequence of launches with no host synchronisation
 * and no allocation, so that the call can be captured in a hipGraph. A plan
 * holds the sizes and the device buffers (sparse grids, the nine spline
 * quantities and their coefficients: 328 B per (knot, mode), plus per-binary
 * tables); plan_bytes returns their total. */
typedef struct bbhx_direct_plan bbhx_direct_plan;

/* lengths_host (B,): sparse grid length of each binary (>= 4, <= 6144).
 * data_length: N dense bins. min_chunk (>= 64): the smallest number of dense
 * bins per work item that bbhx_hip_direct_fused will be called with (sizes the
 * partial-sum buffer). */
int bbhx_hip_direct_plan_create(bbhx_direct_plan** out, int numBinAll, const int* lengths_host, int numModes,
                                int data_length, int min_chunk);
int bbhx_hip_direct_plan_destroy(bbhx_direct_plan* p);
long bbhx_hip_direct_plan_bytes(const bbhx_direct_plan* p);
/* Dense band of each binary from the last bbhx_hip_direct_fused call on this
 * plan: band_host (B, 2) receives [st, et), the dense bins the binary's
 * template covers (waveformbuild.py:139-156). Waits for the device. */
int bbhx_hip_direct_plan_band(const bbhx_direct_plan* p, int* band_host);
/* Arithmetic of later bbhx_hip_direct_fused calls on this plan (no
 * counterpart in the reference, which computes in FP64 only): 0 = FP64
 * (default); 1 = FP32 bulk: the amplitude and transfer-function cubics, the
 * phase factor e^{i phase} and the mode sum in FP32, while the phase and t(f)
 * cubics (and the reduction of the phase modulo 2 pi) and the noise-weighted
 * inner products stay FP64. The first switch to 1 allocates 192 B per (knot,
 * mode) of FP32-bulk segment records (call it outside graph capture). The
 * gradient entry point always computes in FP64. In the FP32-bulk form the
 * dense-band launch uses 3/2 x grid workgroups (its kernel holds 6 wavefronts
 * per SIMD against 4), so a launch shape tuned for the FP64 form serves both. */
int bbhx_hip_direct_plan_set_precision(bbhx_direct_plan* p, int precision);

/* params (11, B) and t_start, t_end (B,) as for bbhx_hip_hdyn_fused; dataFreqs
 * (N,) sorted; dataChannels (3, N) complex = noise weight x data, as
 * Likelihood keeps them (likelihood.py:100-112); noiseFactors (3, N) the
 * weights sqrt(df / S). Outputs (overwritten, not accumulated): d_h, h_h (B,)
 * complex with the reference's factor 4 (Likelihood.cu:357-375); flags (B,)
 * int, which the caller zeroes: bit 0 / bit 1 set when the observation-window
 * search at the low / high end failed for some mode (the reference raises
 * ValueError there, phenomhm.py:530-600). mf_min, mf_max: the default grid
 * Mf range (1e-4, 0.6 in the reference). Launch shape of the persistent
 * dense-band kernel: chunk (>= the plan's min_chunk) dense bins per work item,
 * block = 64, 128 or 256 threads, grid persistent workgroups; the result does
 * not depend on it beyond the order of the sums. */
int bbhx_hip_direct_fused(bbhx_direct_plan* p, double* d_h, double* h_h, int* flags, const double* params,
                          const double* t_start, const double* t_end, const int* ells_host, const int* mms_host,
                          int numModes, const double* dataFreqs, const double* dataChannels,
                          const double* noiseFactors, const double* qnm, int table_len, double dspin,
                          double initial_t_val, double mf_min, double mf_max, const bbhx_response* orbits,
                          int TDItag, int rescaled, int tdi2, int chunk, int block, int grid, void* stream);

/* Gradient of the fused direct log-likelihood with respect to the 11 sampling
 * coordinates (ln M, q, chi1, chi2, d, phi_ref, cos inc, lam, sin beta, psi,
 * t_ref); no counterpart in the reference, which has no derivatives.
 * enable_grad allocates the plan's gradient buffers once (288 B per (knot,
 * mode) and 4.9 KB per (binary, mode)). direct_fused_grad takes the arguments
 * of bbhx_hip_direct_fused plus grad (11, B) (overwritten) and writes d_h, h_h
 * as that call does. The sparse grids and the observation window are held
 * fixed (their dependence on the parameters is not differentiated). The
 * adjoint records are summed with FP64 atomics, so the gradient is not bit for
 * bit reproducible (its value is, to rounding); d_h, h_h are. */
int bbhx_hip_direct_plan_enable_grad(bbhx_direct_plan* p);
int bbhx_hip_direct_fused_grad(bbhx_direct_plan* p, double* d_h, double* h_h, double* grad, int* flags,
                               const double* params, const double* t_start, const double* t_end,
                               const int* ells_host, const int* mms_host, int numModes, const double* dataFreqs,
                               const double* dataChannels, const double* noiseFactors, const double* qnm,
                               int table_len, double dspin, double initial_t_val, double mf_min, double mf_max,
                               const bbhx_response* orbits, int TDItag, int rescaled, int tdi2, int chunk, int block,
                               int grid, void* stream);

/* Heterodyne constants of a reference template (HeterodynedLikelihood.
 * init_heterodyne_info, likelihood.py:499-533; the reference's unused kernel
 * for it is prep_hdyn, Likelihood.cu:238-352). dataFreqs (N,) sorted, already
 * cut to [nodes[0], nodes[K-1]]; dataChannels, h0 (3, N) complex data and
 * reference template there; psd (3, N); nodes (K,) sparse frequencies; df the
 * dense spacing. Writes dataConstants (4, 3, K) complex in the layout hdyn
 * reads (A0, A1, B0, B1; slot j + 1 = segment [nodes[j], nodes[j+1]), the last
 * dense point left out as the reference's loop does, slot 0 zero) and
 * totals (3,) = d_d, h_h, Re d_h over all N points. Deterministic (fixed
 * summation order). */
int bbhx_hip_hdyn_prep(double* dataConstants, double* totals, const double* dataFreqs, const double* dataChannels,
                       const double* h0, const double* psd, int data_length, const double* nodes, int K, double df,
                       void* stream);

/* SSB-frame observation window of each binary, as BBHWaveformFD computes it
 * before the template (waveformbuild.py:477-484, tSSBfromLframe of
 * transform.py:75-105 with t0 = 0; the reference evaluates it with the array
 * module on the device, so its FFI has no entry for it). lam, beta (B,) device
 * arrays (the parameter rows of the fused entry points); tL_start, tL_end the
 * LISA-frame window in seconds (t_obs_start, t_obs_end times YRSID_SI).
 * Writes t_start, t_end (B,) on the device, asynchronously on the stream. */
int bbhx_hip_window_times(double* t_start, double* t_end, const double* lam, const double* beta, int numBinAll,
                          double tL_start, double tL_end, void* stream);

/* Event timing of the fused entry points on the caller's stream, split in two:
 * set-up (records; for the direct form also grids, splines and bands) and
 * evaluation. bbhx_hip_hdyn_fused, bbhx_hip_hdyn_fused_grad and
 * bbhx_hip_direct_fused record it: enable(1) starts recording (and clears earlier records);
 * read() waits for the recorded events and returns the summed milliseconds and
 * the number of calls, then clears. For benchmarks; off by default. */
int bbhx_hip_timing_enable(int on);
int bbhx_hip_timing_read(double* ms_setup, double* ms_eval, int* ncalls);

/* Releases the library's scratch buffers on the current device, retired ones
 * included (synchronises it). Graphs captured before the call must not be
 * replayed after it. */
int bbhx_hip_release_workspace(void);

/* Library version string. */
const char* bbhx_hip_version(void);

#ifdef __cplusplus
}
#endif

#endif /* BBHX_HIP_H */
