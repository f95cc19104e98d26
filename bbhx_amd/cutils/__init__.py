"""The gfx950 backend: ctypes bindings over ``libbbhx_hip.so`` (``include/bbhx_hip.h``).

Mirrors the reference's backend methods (``cutils/__init__.py:13-23``,
``BBHxBackendMethods``): ``waveform_amp_phase_wrap``,
``get_phenomhm_ringdown_frequencies``, ``get_phenomd_ringdown_frequencies``,
``pyFastLISAResponse``, ``interpolate_wrap``, ``InterpTDI_wrap``,
``direct_sum_wrap``, ``hdyn_wrap``, ``direct_like_wrap``, plus ``xp``.
Device arrays are ``torch`` tensors on the ROCm device (``xp = torch``); every
call is queued on ``torch.cuda.current_stream()``.

There is no CPU fallback: if the HIP library or a GPU is missing, constructing
the backend raises ``BackendUnavailableException``.
"""

from __future__ import annotations

import ctypes as C
import dataclasses
import typing
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent.parent / "libbbhx_hip.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_vp = C.c_void_p

STATUS = {0: "ok", 1: "invalid arguments", 2: "kernel launch/runtime error",
          3: "workspace allocation failed (under graph capture: run the call once on the capture stream first)",
          4: "unsupported (l, m) mode",
          5: "batch above the capacity of the allocation-free entry point (BBHX_HIP_NOALLOC_MAX_RECORDS records)"}

# (template, mode) records held by bbhx_hip_hdyn_fused_noalloc (include/bbhx_hip.h)
NOALLOC_MAX_RECORDS = 24576


class BackendUnavailableException(RuntimeError):
    """The requested backend cannot be used on this machine."""


class HipCallError(RuntimeError):
    """A C-ABI entry point returned a nonzero status."""


_lib = None


def load_library():
    """Load ``libbbhx_hip.so`` (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise BackendUnavailableException(
            f"{_LIB_PATH} is not built; run __graft_entry__.build() (or make -C bbhx_amd/csrc)")
    lib = C.CDLL(str(_LIB_PATH))
    sigs = {
        "bbhx_hip_waveform_amp_phase": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int,
                                        _vp, _vp, C.c_int, _vp],
        "bbhx_hip_get_phenomhm_ringdown_frequencies": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, _vp],
        "bbhx_hip_get_phenomd_ringdown_frequencies": [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int] + [_vp] * 8
        + [C.c_double, C.c_int, _vp],
        "bbhx_hip_response_create": [C.POINTER(_vp)],
        "bbhx_hip_response_destroy": [_vp],
        "bbhx_hip_response_add_orbits": [_vp, C.c_double, C.c_int, _vp, _vp, _vp, _ip, _ip, _ip, C.c_double],
        "bbhx_hip_response_compute": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp],
        "bbhx_hip_interpolate": [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp],
        "bbhx_hip_InterpTDI": [_i64p, _vp, _vp, _vp, _vp, _vp, _vp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                               _i64p, _ip, _ip, _vp],
        "bbhx_hip_direct_sum": [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp],
        "bbhx_hip_hdyn": [_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp],
        "bbhx_hip_direct_like": [_vp, _vp, _vp, _vp, _i64p, _ip, _ip, C.c_int, C.c_int, _vp],
        "bbhx_hip_hdyn_fused": [_vp, _vp, _vp, _vp, _vp, C.c_int, _ip, _ip, C.c_int, _vp, _vp, _vp, C.c_int, _vp,
                                C.c_int, C.c_double, C.c_double, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                _vp, C.c_double, _vp],
        "bbhx_hip_hdyn_fused_noalloc": [_vp, _vp, _vp, _vp, _vp, C.c_int, _ip, _ip, C.c_int, _vp, _vp, _vp, C.c_int,
                                        _vp, C.c_int, C.c_double, C.c_double, _vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, _vp, C.c_double, _vp],
        "bbhx_hip_hdyn_fused_grad": [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _ip, _ip, C.c_int, _vp, _vp, _vp,
                                     C.c_int, _vp, C.c_int, C.c_double, C.c_double, _vp, C.c_int, C.c_int, C.c_int,
                                     C.c_int, _vp, C.c_double, _vp],
        "bbhx_hip_direct_plan_create": [C.POINTER(_vp), C.c_int, _ip, C.c_int, C.c_int, C.c_int],
        "bbhx_hip_direct_plan_destroy": [_vp],
        "bbhx_hip_direct_plan_bytes": [_vp],
        "bbhx_hip_direct_plan_band": [_vp, _ip],
        "bbhx_hip_hdyn_prep": [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp, C.c_int, C.c_double, _vp],
        "bbhx_hip_direct_fused": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _ip, _ip, C.c_int, _vp, _vp, _vp, _vp, C.c_int,
                                  C.c_double, C.c_double, C.c_double, C.c_double, _vp, C.c_int, C.c_int, C.c_int,
                                  C.c_int, C.c_int, C.c_int, _vp],
        "bbhx_hip_direct_fused_grad": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _ip, _ip, C.c_int, _vp, _vp, _vp, _vp,
                                       C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _vp, C.c_int,
                                       C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp],
        "bbhx_hip_direct_plan_enable_grad": [_vp],
        "bbhx_hip_direct_plan_set_precision": [_vp, C.c_int],
        "bbhx_hip_window_times": [_vp, _vp, _vp, _vp, C.c_int, C.c_double, C.c_double, _vp],
        "bbhx_hip_release_workspace": [],
        "bbhx_hip_version": [],
        "bbhx_hip_timing_enable": [C.c_int],
        "bbhx_hip_timing_read": [_dp, _dp, _ip],
    }
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = {"bbhx_hip_version": C.c_char_p, "bbhx_hip_direct_plan_bytes": C.c_long}.get(name, C.c_int)
    _lib = lib
    return lib


def exported_symbols():
    """Names of the entry points this module binds (all declared in include/bbhx_hip.h)."""
    return [
        "bbhx_hip_waveform_amp_phase", "bbhx_hip_get_phenomhm_ringdown_frequencies",
        "bbhx_hip_get_phenomd_ringdown_frequencies", "bbhx_hip_response_create", "bbhx_hip_response_destroy",
        "bbhx_hip_response_add_orbits", "bbhx_hip_response_compute", "bbhx_hip_interpolate", "bbhx_hip_InterpTDI",
        "bbhx_hip_direct_sum", "bbhx_hip_hdyn", "bbhx_hip_direct_like", "bbhx_hip_hdyn_fused",
        "bbhx_hip_hdyn_fused_noalloc",
        "bbhx_hip_hdyn_fused_grad", "bbhx_hip_direct_plan_create", "bbhx_hip_direct_plan_destroy",
        "bbhx_hip_direct_plan_bytes", "bbhx_hip_direct_plan_band", "bbhx_hip_direct_fused", "bbhx_hip_hdyn_prep",
        "bbhx_hip_direct_fused_grad", "bbhx_hip_direct_plan_enable_grad", "bbhx_hip_direct_plan_set_precision",
        "bbhx_hip_window_times", "bbhx_hip_release_workspace", "bbhx_hip_version", "bbhx_hip_timing_enable",
        "bbhx_hip_timing_read",
    ]


def _check(status, name):
    if status != 0:
        raise HipCallError(f"{name} failed with status {status} ({STATUS.get(status, 'unknown')})")


# ---------------------------------------------------------------- helpers


def _torch():
    import torch

    return torch


def _stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _dev(t, dtype=None, name="array"):
    """Device pointer of a contiguous CUDA tensor of the given dtype."""
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor on the GPU, got {type(t)}")
    if not t.is_cuda:
        raise ValueError(f"{name} must be on the GPU")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{name} must have dtype {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return C.c_void_p(t.data_ptr())


def _host_i64(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int64))
    return a, a.ctypes.data_as(_i64p)


def _host_i32(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32))
    return a, a.ctypes.data_as(_ip)


def _host_f64(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    return a, a.ctypes.data_as(_dp)


# ---------------------------------------------------------------- methods


def waveform_amp_phase_wrap(waveformOut, ells, mms, freqs, m1_SI, m2_SI, chi1z, chi2z, distance, f_ref,
                            numModes, length, numBinAll, Mf_RD_lm_all, Mf_DM_lm_all, run_phenomd):
    """cutils/phenomhm.pyx:61 -> PhenomHMWaveform.cu:2998."""
    torch = _torch()
    f64, i32 = torch.float64, torch.int32
    st = load_library().bbhx_hip_waveform_amp_phase(
        _dev(waveformOut, f64, "waveformOut"), _dev(ells, i32, "ells"), _dev(mms, i32, "mms"),
        _dev(freqs, f64, "freqs"), _dev(m1_SI, f64), _dev(m2_SI, f64), _dev(chi1z, f64), _dev(chi2z, f64),
        _dev(distance, f64), _dev(f_ref, f64), int(numModes), int(length), int(numBinAll),
        _dev(Mf_RD_lm_all, f64), _dev(Mf_DM_lm_all, f64), int(bool(run_phenomd)), _stream())
    _check(st, "waveform_amp_phase")


def get_phenomhm_ringdown_frequencies(fringdown, fdamp, m1, m2, chi1z, chi2z, ells, mms, numModes, numBinAll):
    """cutils/phenomhm.pyx:117 -> PhenomHMWaveform.cu:1879."""
    torch = _torch()
    f64, i32 = torch.float64, torch.int32
    st = load_library().bbhx_hip_get_phenomhm_ringdown_frequencies(
        _dev(fringdown, f64), _dev(fdamp, f64), _dev(m1, f64), _dev(m2, f64), _dev(chi1z, f64), _dev(chi2z, f64),
        _dev(ells, i32), _dev(mms, i32), int(numModes), int(numBinAll), _stream())
    _check(st, "get_phenomhm_ringdown_frequencies")


def get_phenomd_ringdown_frequencies(fringdown, fdamp, m1, m2, chi1z, chi2z, numBinAll, y_rd, c1_rd, c2_rd, c3_rd,
                                     y_dm, c1_dm, c2_dm, c3_dm, dspin):
    """cutils/phenomhm.pyx:155 -> PhenomHMWaveform.cu:1995."""
    torch = _torch()
    f64 = torch.float64
    tabs = [y_rd, c1_rd, c2_rd, c3_rd, y_dm, c1_dm, c2_dm, c3_dm]
    n = min(int(t.numel()) for t in tabs)
    st = load_library().bbhx_hip_get_phenomd_ringdown_frequencies(
        _dev(fringdown, f64), _dev(fdamp, f64), _dev(m1, f64), _dev(m2, f64), _dev(chi1z, f64), _dev(chi2z, f64),
        int(numBinAll), *[_dev(t, f64) for t in tabs], float(dspin), n, _stream())
    _check(st, "get_phenomd_ringdown_frequencies")


class pyFastLISAResponse:
    """Stateful response object (cutils/lisaresponse.pyx:34-121)."""

    def __init__(self):
        lib = load_library()
        h = C.c_void_p()
        _check(lib.bbhx_hip_response_create(C.byref(h)), "response_create")
        self._h = h
        self._keep = ()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.bbhx_hip_response_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def add_orbit_information(self, dt, N, n_arr, L_arr, x_arr, links, sc_r, sc_e, armlength):
        """lisaresponse.pyx:41-71. n_arr, L_arr, x_arr: device float64 tensors; links/sc_r/sc_e: 6 ints."""
        torch = _torch()
        links_h, links_p = _host_i32(links)
        scr_h, scr_p = _host_i32(sc_r)
        sce_h, sce_p = _host_i32(sc_e)
        st = load_library().bbhx_hip_response_add_orbits(
            self._h, float(dt), int(N), _dev(n_arr, torch.float64, "n_arr"), _dev(L_arr, torch.float64, "L_arr"),
            _dev(x_arr, torch.float64, "x_arr"), links_p, scr_p, sce_p, float(armlength))
        _check(st, "response_add_orbits")
        # the library keeps the device pointers: hold the tensors
        self._keep = (n_arr, L_arr, x_arr)

    def LISA_response_wrap(self, response_out, ells, mms, freqs, phi_ref, inc, lam, beta, psi, TDItag, rescaled,
                           tdi2, order_fresnel_stencil, numModes, length, numBinAll, includesAmps):
        """lisaresponse.pyx:78 -> Response.cu:671."""
        torch = _torch()
        f64, i32 = torch.float64, torch.int32
        st = load_library().bbhx_hip_response_compute(
            self._h, _dev(response_out, f64, "response_out"), _dev(ells, i32), _dev(mms, i32), _dev(freqs, f64),
            _dev(phi_ref, f64), _dev(inc, f64), _dev(lam, f64), _dev(beta, f64), _dev(psi, f64), int(TDItag),
            int(bool(rescaled)), int(bool(tdi2)), int(order_fresnel_stencil), int(numModes), int(length),
            int(numBinAll), int(includesAmps), _stream())
        _check(st, "LISA_response")


def interpolate_wrap(freqs, propArrays, B, upper_diag, diag, lower_diag, length, numInterpParams, numModes,
                     numBinAll):
    """cutils/interp.pyx:14 -> Interpolate.cu:320."""
    torch = _torch()
    f64 = torch.float64
    st = load_library().bbhx_hip_interpolate(
        _dev(freqs, f64), _dev(propArrays, f64), _dev(B, f64), _dev(upper_diag, f64), _dev(diag, f64),
        _dev(lower_diag, f64), int(length), int(numInterpParams), int(numModes), int(numBinAll), _stream())
    _check(st, "interpolate")


def InterpTDI_wrap(templateChannels_ptrs, dataFreqs, freqs, propArrays, c1, c2, c3, t_start, t_end, length,
                   data_length, numBinAll, numModes, inds_ptrs, inds_start, ind_lengths):
    """cutils/bbhwaveformbuild.pyx:18 -> WaveformBuild.cu:147. Pointer arrays, starts,
    lengths and t_start/t_end are host arrays (as in the reference)."""
    torch = _torch()
    f64 = torch.float64
    tp, tpp = _host_i64(templateChannels_ptrs)
    ip, ipp = _host_i64(inds_ptrs)
    s, sp = _host_i32(inds_start)
    le, lp = _host_i32(ind_lengths)
    ts, tsp = _host_f64(np.atleast_1d(t_start))
    te, tep = _host_f64(np.atleast_1d(t_end))
    st = load_library().bbhx_hip_InterpTDI(
        tpp, _dev(dataFreqs, f64), _dev(freqs, f64), _dev(propArrays, f64), _dev(c1, f64), _dev(c2, f64),
        _dev(c3, f64), tsp, tep, int(length), int(data_length), int(numBinAll), int(numModes), ipp, sp, lp,
        _stream())
    _check(st, "InterpTDI")


def direct_sum_wrap(templateChannels, bbh_buffer, numBinAll, data_length, nChannels, numModes, t_start, t_end):
    """cutils/bbhwaveformbuild.pyx:37 -> WaveformBuild.cu:281."""
    torch = _torch()
    st = load_library().bbhx_hip_direct_sum(
        _dev(templateChannels, torch.complex128), _dev(bbh_buffer, torch.float64), int(numBinAll), int(data_length),
        int(nChannels), int(numModes), _dev(t_start, torch.float64), _dev(t_end, torch.float64), _stream())
    _check(st, "direct_sum")


def hdyn_wrap(likeOut1, likeOut2, templateChannels, dataConstants, dataFreqs, numBinAll, data_length, nChannels,
              include_last_segment=True):
    """cutils/bbhlikelihood.pyx:21 -> Likelihood.cu:187."""
    torch = _torch()
    c128 = torch.complex128
    st = load_library().bbhx_hip_hdyn(
        _dev(likeOut1, c128), _dev(likeOut2, c128), _dev(templateChannels, c128), _dev(dataConstants, c128),
        _dev(dataFreqs, torch.float64), int(numBinAll), int(data_length), int(nChannels),
        int(bool(include_last_segment)), _stream())
    _check(st, "hdyn")


def direct_like_wrap(d_h, h_h, dataChannels, noise_weight_times_df, templateChannels_ptrs, inds_start, ind_lengths,
                     data_stream_length, numBinAll):
    """cutils/bbhlikelihood.pyx:42 -> Likelihood.cu:382. d_h, h_h: device complex tensors (+=)."""
    torch = _torch()
    c128 = torch.complex128
    tp, tpp = _host_i64(templateChannels_ptrs)
    s, sp = _host_i32(inds_start)
    le, lp = _host_i32(ind_lengths)
    st = load_library().bbhx_hip_direct_like(
        _dev(d_h, c128), _dev(h_h, c128), _dev(dataChannels, c128), _dev(noise_weight_times_df, torch.float64), tpp,
        sp, lp, int(data_stream_length), int(numBinAll), _stream())
    _check(st, "direct_like")


def _need(t, n, name):
    if int(t.numel()) != int(n):
        raise ValueError(f"{name} must hold {int(n)} elements, got {int(t.numel())}")


def _check_fused_sizes(params, B, d_h, h_h, t_start, t_end, logL, qnm):
    """Host-side size checks shared by the fused entry points (before any device call)."""
    if params.dim() != 2 or int(params.shape[0]) != 11:
        raise ValueError(f"params must be (11, B), got {tuple(params.shape)}")
    for t, name in ((d_h, "d_h"), (h_h, "h_h"), (t_start, "t_start"), (t_end, "t_end")):
        _need(t, B, name)
    if logL is not None:
        _need(logL, B, "logL")
    if int(qnm.numel()) < 16 or int(qnm.numel()) % 8:
        raise ValueError("qnm must hold the 8 rows of the QNM table")


def _check_hdyn_sizes(params, d_h, h_h, t_start, t_end, freqs, h0, data_constants, qnm, logL):
    B, K = int(params.shape[1]) if params.dim() == 2 else -1, int(freqs.numel())
    _check_fused_sizes(params, B, d_h, h_h, t_start, t_end, logL, qnm)
    _need(h0, 3 * K, "h0 (3, K)")
    _need(data_constants, 12 * K, "data_constants (4, 3, K)")
    return B, K


def hdyn_fused(d_h, h_h, params, t_start, t_end, ells, mms, freqs, h0, data_constants, qnm, dspin, initial_t_val,
               response, TDItag, rescaled, tdi2, include_last_segment=True, precision=0, logL=None, d_d=0.0,
               noalloc=False):
    """Fused heterodyned likelihood (no reference counterpart; see include/bbhx_hip.h).
    ``logL`` (B,) float64 device, optional: also written, -1/2 (d_d + Re h_h - 2 Re d_h).
    ``noalloc``: run bbhx_hip_hdyn_fused_noalloc (no device allocation; at most
    NOALLOC_MAX_RECORDS (template, mode) records; calls on one GPU must not overlap)."""
    torch = _torch()
    f64, c128 = torch.float64, torch.complex128
    B, K = _check_hdyn_sizes(params, d_h, h_h, t_start, t_end, freqs, h0, data_constants, qnm, logL)
    el, elp = _host_i32(ells)
    mm, mmp = _host_i32(mms)
    table_len = int(qnm.numel()) // 8
    lib = load_library()
    entry = lib.bbhx_hip_hdyn_fused_noalloc if noalloc else lib.bbhx_hip_hdyn_fused
    st = entry(
        _dev(d_h, c128), _dev(h_h, c128), _dev(params, f64, "params"), _dev(t_start, f64), _dev(t_end, f64), B,
        elp, mmp, int(el.size), _dev(freqs, f64), _dev(h0, c128, "h0"), _dev(data_constants, c128), K,
        _dev(qnm, f64), table_len, float(dspin), float(initial_t_val), response.handle, int(TDItag),
        int(bool(rescaled)), int(bool(tdi2)), int(bool(include_last_segment)), int(precision),
        None if logL is None else _dev(logL, f64, "logL"), float(d_d), _stream())
    _check(st, "hdyn_fused_noalloc" if noalloc else "hdyn_fused")


def hdyn_fused_grad(d_h, h_h, grad, params, t_start, t_end, ells, mms, freqs, h0, data_constants, qnm, dspin,
                    initial_t_val, response, TDItag, rescaled, tdi2, include_last_segment=True, logL=None, d_d=0.0):
    """Heterodyned likelihood and its gradient (11, B) (no reference counterpart; see include/bbhx_hip.h).
    ``logL`` (B,) float64 device, optional: also written, as by ``hdyn_fused``."""
    torch = _torch()
    f64, c128 = torch.float64, torch.complex128
    B, K = _check_hdyn_sizes(params, d_h, h_h, t_start, t_end, freqs, h0, data_constants, qnm, logL)
    el, elp = _host_i32(ells)
    mm, mmp = _host_i32(mms)
    table_len = int(qnm.numel()) // 8
    if tuple(grad.shape) != (11, B):
        raise ValueError("grad must be (11, B)")
    st = load_library().bbhx_hip_hdyn_fused_grad(
        _dev(d_h, c128), _dev(h_h, c128), _dev(grad, f64, "grad"), _dev(params, f64, "params"), _dev(t_start, f64),
        _dev(t_end, f64), B, elp, mmp, int(el.size), _dev(freqs, f64), _dev(h0, c128, "h0"),
        _dev(data_constants, c128), K, _dev(qnm, f64), table_len, float(dspin), float(initial_t_val),
        response.handle, int(TDItag), int(bool(rescaled)), int(bool(tdi2)), int(bool(include_last_segment)),
        None if logL is None else _dev(logL, f64, "logL"), float(d_d), _stream())
    _check(st, "hdyn_fused_grad")


class DirectPlan:
    """Sizes and device buffers of the fused direct likelihood of a ragged batch
    (bbhx_hip_direct_plan_*). lengths: (B,) sparse grid lengths. The launch
    shape (chunk, block, grid) of the dense-band kernel can change between
    calls (set_launch; chunk >= min_chunk)."""

    def __init__(self, lengths, num_modes, data_length, chunk=16384, block=256, grid=None, min_chunk=1024):
        torch = _torch()
        lib = load_library()
        lens, lp = _host_i32(lengths)
        if grid is None:
            grid = 8 * torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
        self.lengths = lens.copy()
        self.num_modes, self.data_length = int(num_modes), int(data_length)
        self.min_chunk = int(min(min_chunk, chunk))
        self.set_launch(chunk, block, grid)
        self.precision = 0
        h = _vp()
        _check(lib.bbhx_hip_direct_plan_create(C.byref(h), int(lens.size), lp, self.num_modes, self.data_length,
                                               self.min_chunk), "direct_plan_create")
        self._h = h

    def set_launch(self, chunk, block, grid):
        if int(chunk) < self.min_chunk:
            raise ValueError(f"chunk must be >= {self.min_chunk} for this plan")
        self.chunk, self.block, self.grid = int(chunk), int(block), int(grid)

    @property
    def handle(self):
        return self._h

    @property
    def num_bin_all(self):
        return int(self.lengths.size)

    def nbytes(self):
        return int(load_library().bbhx_hip_direct_plan_bytes(self._h))

    def enable_grad(self):
        """Allocate the gradient buffers (288 B per (knot, mode) + records with tangents)."""
        _check(load_library().bbhx_hip_direct_plan_enable_grad(self._h), "direct_plan_enable_grad")

    def set_precision(self, precision):
        """0: FP64; 1: FP32 bulk (amplitude and transfer-function cubics, phase
        factor and mode sum in FP32; phase, t(f) and the inner products in FP64).
        The first switch to 1 allocates 192 B per (knot, mode)."""
        precision = int(precision)
        if precision != self.precision:
            _check(load_library().bbhx_hip_direct_plan_set_precision(self._h, precision), "direct_plan_set_precision")
            self.precision = precision

    def band(self):
        """(B, 2) int32 dense band [st, et) of each binary from the last call."""
        out = np.zeros((self.num_bin_all, 2), dtype=np.int32)
        _check(load_library().bbhx_hip_direct_plan_band(self._h, out.ctypes.data_as(_ip)), "direct_plan_band")
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                load_library().bbhx_hip_direct_plan_destroy(h)
            except Exception:  # pragma: no cover - interpreter shutdown
                pass
            self._h = None


def direct_fused(plan, d_h, h_h, flags, params, t_start, t_end, ells, mms, data_freqs, data_channels, noise_factors,
                 qnm, dspin, initial_t_val, mf_min, mf_max, response, TDItag, rescaled, tdi2):
    """Fused direct likelihood of a ragged batch (no reference counterpart; see include/bbhx_hip.h)."""
    torch = _torch()
    f64, c128 = torch.float64, torch.complex128
    el, elp = _host_i32(ells)
    mm, mmp = _host_i32(mms)
    B = plan.num_bin_all
    if tuple(params.shape) != (11, B):
        raise ValueError("params must be (11, B) for this plan")
    _check_fused_sizes(params, B, d_h, h_h, t_start, t_end, None, qnm)
    _need(flags, B, "flags")
    N = plan.data_length
    _need(data_freqs, N, "data_freqs")
    _need(data_channels, 3 * N, "data_channels (3, N)")
    _need(noise_factors, 3 * N, "noise_factors (3, N)")
    st = load_library().bbhx_hip_direct_fused(
        plan.handle, _dev(d_h, c128), _dev(h_h, c128), _dev(flags, torch.int32, "flags"), _dev(params, f64, "params"),
        _dev(t_start, f64), _dev(t_end, f64), elp, mmp, int(el.size), _dev(data_freqs, f64),
        _dev(data_channels, c128), _dev(noise_factors, f64), _dev(qnm, f64), int(qnm.numel()) // 8, float(dspin),
        float(initial_t_val), float(mf_min), float(mf_max), response.handle, int(TDItag), int(bool(rescaled)),
        int(bool(tdi2)), plan.chunk, plan.block, plan.grid, _stream())
    _check(st, "direct_fused")


def direct_fused_grad(plan, d_h, h_h, grad, flags, params, t_start, t_end, ells, mms, data_freqs, data_channels,
                      noise_factors, qnm, dspin, initial_t_val, mf_min, mf_max, response, TDItag, rescaled, tdi2):
    """Fused direct likelihood of a ragged batch and its (11, B) gradient (no reference counterpart)."""
    torch = _torch()
    f64, c128 = torch.float64, torch.complex128
    el, elp = _host_i32(ells)
    mm, mmp = _host_i32(mms)
    B = plan.num_bin_all
    if tuple(params.shape) != (11, B):
        raise ValueError("params must be (11, B) for this plan")
    _check_fused_sizes(params, B, d_h, h_h, t_start, t_end, None, qnm)
    _need(flags, B, "flags")
    N = plan.data_length
    _need(data_freqs, N, "data_freqs")
    _need(data_channels, 3 * N, "data_channels (3, N)")
    _need(noise_factors, 3 * N, "noise_factors (3, N)")
    st = load_library().bbhx_hip_direct_fused_grad(
        plan.handle, _dev(d_h, c128), _dev(h_h, c128), _dev(grad, f64, "grad"), _dev(flags, torch.int32, "flags"),
        _dev(params, f64, "params"),
        _dev(t_start, f64), _dev(t_end, f64), elp, mmp, int(el.size), _dev(data_freqs, f64),
        _dev(data_channels, c128), _dev(noise_factors, f64), _dev(qnm, f64), int(qnm.numel()) // 8, float(dspin),
        float(initial_t_val), float(mf_min), float(mf_max), response.handle, int(TDItag), int(bool(rescaled)),
        int(bool(tdi2)), plan.chunk, plan.block, plan.grid, _stream())
    _check(st, "direct_fused_grad")


def hdyn_prep(data_constants, totals, data_freqs, data_channels, h0, psd, nodes, df):
    """Heterodyne constants on the device (no reference counterpart in BBHxBackendMethods; see include/bbhx_hip.h)."""
    torch = _torch()
    f64, c128 = torch.float64, torch.complex128
    N, K = int(data_freqs.numel()), int(nodes.numel())
    if tuple(data_channels.shape) != (3, N) or tuple(h0.shape) != (3, N) or tuple(psd.shape) != (3, N):
        raise ValueError("data_channels, h0 and psd must be (3, N)")
    if data_constants.numel() != 12 * K or totals.numel() != 3:
        raise ValueError("data_constants must hold 4 x 3 x K complex, totals 3 doubles")
    st = load_library().bbhx_hip_hdyn_prep(_dev(data_constants, c128), _dev(totals, f64), _dev(data_freqs, f64),
                                           _dev(data_channels, c128), _dev(h0, c128), _dev(psd, f64), N,
                                           _dev(nodes, f64), K, float(df), _stream())
    _check(st, "hdyn_prep")


def window_times(t_start, t_end, lam, beta, tL_start, tL_end):
    """SSB-frame window (t_start, t_end) of each binary from device rows lam,
    beta (tSSBfromLframe at the LISA-frame times tL_start, tL_end in seconds),
    written on the device on the current stream (no reference counterpart in
    BBHxBackendMethods; see include/bbhx_hip.h)."""
    torch = _torch()
    f64 = torch.float64
    B = int(lam.numel())
    if beta.numel() != B or t_start.numel() != B or t_end.numel() != B:
        raise ValueError("lam, beta, t_start and t_end must have one entry per binary")
    st = load_library().bbhx_hip_window_times(_dev(t_start, f64), _dev(t_end, f64), _dev(lam, f64), _dev(beta, f64),
                                              B, float(tL_start), float(tL_end), _stream())
    _check(st, "window_times")


def timing_enable(on=True):
    """Start (or stop) event timing of the fused heterodyne launches."""
    _check(load_library().bbhx_hip_timing_enable(int(bool(on))), "timing_enable")


def timing_read():
    """(ms_setup, ms_eval, ncalls) summed over the calls since timing_enable."""
    a, b, n = C.c_double(), C.c_double(), C.c_int()
    _check(load_library().bbhx_hip_timing_read(C.byref(a), C.byref(b), C.byref(n)), "timing_read")
    return a.value, b.value, n.value


# ---------------------------------------------------------------- backend


@dataclasses.dataclass
class BBHxBackendMethods:
    """The reference's method set (cutils/__init__.py:13-23)."""

    hdyn_wrap: typing.Callable[..., None]
    direct_like_wrap: typing.Callable[..., None]
    direct_sum_wrap: typing.Callable[..., None]
    InterpTDI_wrap: typing.Callable[..., None]
    pyFastLISAResponse: object
    waveform_amp_phase_wrap: typing.Callable[..., None]
    get_phenomhm_ringdown_frequencies: typing.Callable[..., None]
    get_phenomd_ringdown_frequencies: typing.Callable[..., None]
    interpolate_wrap: typing.Callable[..., None]
    xp: object


class BBHxHipBackend:
    """The gfx950 backend ("bbhx_hip"). Raises BackendUnavailableException when
    the GPU or the HIP library is missing; there is no fallback."""

    _name = "bbhx_hip"
    uses_gpu = True

    def __init__(self):
        torch = _torch()
        if not torch.cuda.is_available():
            raise BackendUnavailableException("'hip' backend needs a ROCm GPU (torch.cuda.is_available() is False)")
        load_library()
        m = BBHxBackendMethods(
            hdyn_wrap=hdyn_wrap,
            direct_like_wrap=direct_like_wrap,
            direct_sum_wrap=direct_sum_wrap,
            InterpTDI_wrap=InterpTDI_wrap,
            pyFastLISAResponse=pyFastLISAResponse,
            waveform_amp_phase_wrap=waveform_amp_phase_wrap,
            get_phenomhm_ringdown_frequencies=get_phenomhm_ringdown_frequencies,
            get_phenomd_ringdown_frequencies=get_phenomd_ringdown_frequencies,
            interpolate_wrap=interpolate_wrap,
            xp=torch,
        )
        for f in dataclasses.fields(m):
            setattr(self, f.name, getattr(m, f.name))
        self.hdyn_fused = hdyn_fused
        self.hdyn_fused_grad = hdyn_fused_grad
        self.direct_fused = direct_fused
        self.DirectPlan = DirectPlan
        self.hdyn_prep = hdyn_prep
        self.window_times = window_times
        self.direct_fused_grad = direct_fused_grad
        self.device = torch.device("cuda", torch.cuda.current_device())

    @property
    def name(self):
        return self._name


__all__ = ["BBHxHipBackend", "BBHxBackendMethods", "BackendUnavailableException", "HipCallError", "load_library",
           "exported_symbols"]
