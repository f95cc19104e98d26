// LISA TDI response kernels and the stateful response object's C-ABI
// (the counterpart of the reference's pyFastLISAResponse, lisaresponse.pyx:34-121,
// and FastLISAResponse::LISA_response, Response.cu:671-746).
//
//  k_resp_setup  one lane per (binary, mode): H matrix and wave vector
//  k_resp_eval   one lane per (binary, mode, frequency): transfer functions,
//                in-place phase += orbital Doppler phase

#include <cstdlib>
#include <new>

#include "bbhx_hip.h"
#include "common.h"
#include "response_device.h"
#include "workspace.h"

using namespace bbhx;
using namespace bbhx::resp;

struct bbhx_response {
    const double* x = nullptr;  // (N, 3, 3) device
    const double* n = nullptr;  // (N, 6, 3) device
    double* packed = nullptr;   // (N, ORB_ROW) device, owned: packed rows of x and n
    double dt = 0.0;
    int N = 0;
    int li12 = -1, li23 = -1, li31 = -1;
    double armlength = 0.0;
};

namespace {

constexpr int EVAL_THREADS = 256;

__global__ __launch_bounds__(64) void k_resp_setup(RespRec* __restrict__ recs, const int* __restrict__ ells, const int* __restrict__ mms,
                             const double* __restrict__ phi_ref, const double* __restrict__ inc,
                             const double* __restrict__ lam, const double* __restrict__ beta,
                             const double* __restrict__ psi, int numModes, int numBinAll)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= numBinAll * numModes) return;
    int b = idx / numModes, mi = idx % numModes;
    RespRec r;
    resp_setup(r, ells[mi], mms[mi], phi_ref[b], inc[b], lam[b], beta[b], psi[b]);
    recs[idx] = r;
}

__global__ __launch_bounds__(EVAL_THREADS) void k_resp_eval(double* __restrict__ phases, const double* __restrict__ tf,
                                                            double* __restrict__ resp,
                                                            const RespRec* __restrict__ recs,
                                                            const double* __restrict__ freqs, OrbitTab orb,
                                                            int TDItag, int rescaled, int tdi2, int numModes,
                                                            int length, long nper)
{
    const int i = blockIdx.x * EVAL_THREADS + threadIdx.x;
    const int mi = blockIdx.y, b = blockIdx.z;
    if (i >= length) return;
    const RespRec& r = recs[b * numModes + mi];
    const long idx = ((long)b * numModes + mi) * length + i;
    cd T[3];
    double phaseRdelay;
    resp_eval(r, orb, freqs[idx], tf[idx], TDItag, rescaled, tdi2, T, phaseRdelay);
    resp[0 * nper + idx] = T[0].re;
    resp[1 * nper + idx] = T[0].im;
    resp[2 * nper + idx] = T[1].re;
    resp[3 * nper + idx] = T[1].im;
    resp[4 * nper + idx] = T[2].re;
    resp[5 * nper + idx] = T[2].im;
    phases[idx] += phaseRdelay;
}

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Packed orbit rows (OrbitTab::p): one thread per (row, value).
__global__ __launch_bounds__(256) void k_orbit_pack(double* __restrict__ p, const double* __restrict__ x,
                                                    const double* __restrict__ n, int N, int li12, int li23, int li31)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)N * ORB_VALS) return;
    const int w = (int)(i / ORB_VALS), q = (int)(i % ORB_VALS);
    auto val = [&](int row) {
        if (q < 9) return x[(long)row * 9 + q];
        const int j = (q - 9) / 3, c = (q - 9) % 3;
        const int li = j == 0 ? li12 : (j == 1 ? li23 : li31);
        return n[(long)row * 18 + 3 * li + c];
    };
    const double v = val(w);
    p[(long)w * ORB_ROW + q] = v;
    p[(long)w * ORB_ROW + ORB_VALS + q] = w + 1 < N ? val(w + 1) - v : 0.0;
}

}  // namespace

namespace bbhx {
// Used by the fused kernels: orbit table view of a response object.
bool response_orbit_tab(const bbhx_response* r, OrbitTab& o)
{
    if (!r || !r->x || !r->n || !r->packed || r->N < 2 || r->li12 < 0 || r->li23 < 0 || r->li31 < 0) return false;
    o.x = r->x;
    o.n = r->n;
    o.p = r->packed;
    o.dt = r->dt;
    o.inv_dt = 1.0 / r->dt;
    o.N = r->N;
    o.li12 = r->li12;
    o.li23 = r->li23;
    o.li31 = r->li31;
    o.L = r->armlength;
    o.inv_L = 1.0 / r->armlength;
    return true;
}
}  // namespace bbhx

extern "C" {

int bbhx_hip_response_create(bbhx_response** out)
{
    if (!out) return ERR_ARGS;
    *out = new (std::nothrow) bbhx_response();
    return *out ? OK : ERR_ALLOC;
}

int bbhx_hip_response_destroy(bbhx_response* r)
{
    if (r && r->packed) {
        (void)hipDeviceSynchronize();
        (void)hipFree(r->packed);
    }
    delete r;
    return OK;
}

int bbhx_hip_response_add_orbits(bbhx_response* r, double dt, int N, const double* n_arr, const double* L_arr,
                                 const double* x_arr, const int* links, const int* sc_r, const int* sc_e,
                                 double armlength)
{
    (void)L_arr;  // light travel times are not used by the equal-arm response
    (void)sc_r;
    (void)sc_e;
    if (!r || !(dt > 0.0) || N < 2 || !n_arr || !x_arr || !links || !(armlength > 0.0)) return ERR_ARGS;
    int li12 = -1, li23 = -1, li31 = -1;
    for (int j = 0; j < 6; j++) {
        if (links[j] == 12) li12 = j;
        if (links[j] == 23) li23 = j;
        if (links[j] == 31) li31 = j;
    }
    if (li12 < 0 || li23 < 0 || li31 < 0) return ERR_ARGS;
    // packed rows, built once here (device-wide synchronisation on both sides:
    // the caller's arrays may come from any stream, and kernels of an earlier
    // table may still read the old rows)
    if (hipDeviceSynchronize() != hipSuccess) return ERR_LAUNCH;
    if (r->packed) {
        (void)hipFree(r->packed);
        r->packed = nullptr;
    }
    double* packed = nullptr;
    if (hipMalloc(&packed, sizeof(double) * ORB_ROW * (size_t)N) != hipSuccess) return ERR_ALLOC;
    hipLaunchKernelGGL(k_orbit_pack, dim3(ceil_div(N * ORB_VALS, 256)), dim3(256), 0, nullptr, packed, x_arr, n_arr, N,
                       li12, li23, li31);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(packed);
        return ERR_LAUNCH;
    }
    r->packed = packed;
    r->x = x_arr;
    r->n = n_arr;
    r->dt = dt;
    r->N = N;
    r->li12 = li12;
    r->li23 = li23;
    r->li31 = li31;
    r->armlength = armlength;
    return OK;
}

int bbhx_hip_response_compute(bbhx_response* r, double* response_out, const int* ells, const int* mms,
                              const double* freqs, const double* phi_ref, const double* inc, const double* lam,
                              const double* beta, const double* psi, int TDItag, int rescaled, int tdi2,
                              int order_fresnel_stencil, int numModes, int length, int numBinAll, int includesAmps,
                              void* stream)
{
    OrbitTab orb;
    if (!response_orbit_tab(r, orb)) return ERR_ARGS;
    if (order_fresnel_stencil != 0) return ERR_ARGS;
    if (numBinAll < 0 || length < 1 || numModes < 1 || numModes > MAX_MODES) return ERR_ARGS;
    if (includesAmps != 0 && includesAmps != 1) return ERR_ARGS;
    if (numBinAll == 0) return OK;
    if (!response_out || !ells || !mms || !freqs || !phi_ref || !inc || !lam || !beta || !psi) return ERR_ARGS;
    const long nper = (long)numBinAll * numModes * length;
    double* phases = response_out + includesAmps * nper;
    const double* tf = response_out + (includesAmps + 1) * nper;
    double* resp = response_out + (includesAmps + 2) * nper;
    const int nrec = numBinAll * numModes;
    RespRec* recs = static_cast<RespRec*>(workspace_get(WS_RESP_REC, sizeof(RespRec) * (size_t)nrec, S(stream)));
    if (!recs) return ERR_ALLOC;
    hipLaunchKernelGGL(k_resp_setup, dim3(ceil_div(nrec, 64)), dim3(64), 0, S(stream), recs, ells, mms, phi_ref, inc,
                       lam, beta, psi, numModes, numBinAll);
    if (hipGetLastError() != hipSuccess) return ERR_LAUNCH;
    for (int b0 = 0; b0 < numBinAll; b0 += 65535) {
        int nb = numBinAll - b0 < 65535 ? numBinAll - b0 : 65535;
        long off = (long)b0 * numModes * length;
        hipLaunchKernelGGL(k_resp_eval, dim3(ceil_div(length, EVAL_THREADS), numModes, nb), dim3(EVAL_THREADS), 0,
                           S(stream), phases + off, tf + off, resp + off, recs + (long)b0 * numModes, freqs + off, orb,
                           TDItag, rescaled, tdi2, numModes, length, nper);
    }
    return hipGetLastError() == hipSuccess ? OK : ERR_LAUNCH;
}

}  // extern "C"
