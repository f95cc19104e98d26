"""FP64 CPU oracle for the PhenomHM -> TDI response -> spline -> likelihood path.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` import this package, and only as the checker
or the timed CPU baseline. The shipped package ``bbhx_amd`` never imports it.

Two native libraries sit behind this module:

* ``oracle/liboracle.so``: this project's C++ restatement of the reference's
  CPU algorithm (``oracle_*.cpp``, built by ``oracle/Makefile``);
* ``oracle/_ref/libbbhx_ref.so``: the reference's own CPU code for
  ``cutils/PhenomHMWaveform.cu``, ``WaveformBuild.cu`` and ``Likelihood.cu``,
  compiled from the reference sources by ``make -C oracle ref`` where the
  reference tree is available. It pins the restatement and is optional at run
  time (``ref_available()``).

All arrays are numpy, C-contiguous, float64 / complex128 / int32 as the
reference's flat layouts require.
"""

from __future__ import annotations

import contextlib
import ctypes as C
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liboracle.so"
_REF_PATH = _HERE / "_ref" / "libbbhx_ref.so"
_FMA_PATH = _HERE / "liboracle_fma.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)

_lib = None
_ref = None


def _load():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(f"{_LIB_PATH} is not built; run `make -C oracle` (or __graft_entry__.build())")
        _lib = C.CDLL(str(_LIB_PATH))
    return _lib


@contextlib.contextmanager
def fma_rounding():
    """Inside the block the oracle runs from ``liboracle_fma.so``: the same
    sources compiled with every multiply-add contracted into an FMA. The two
    builds compute the same algorithm with different FP64 rounding, so their
    difference is the spread that rounding alone gives a result (the floor
    below which two FP64 implementations cannot be expected to agree)."""
    global _lib
    saved = _load()
    if not _FMA_PATH.exists():
        raise RuntimeError(f"{_FMA_PATH} is not built; run `make -C oracle`")
    _lib = C.CDLL(str(_FMA_PATH))
    try:
        yield
    finally:
        _lib = saved


def ref_available() -> bool:
    return _REF_PATH.exists()


def _load_ref():
    global _ref
    if _ref is None:
        if not _REF_PATH.exists():
            raise RuntimeError(f"{_REF_PATH} is not built; run `make -C oracle ref` where the reference is available")
        _ref = C.CDLL(str(_REF_PATH))
    return _ref


def _d(a):
    assert a.dtype in (np.float64, np.complex128) and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def _l(a):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_lp)


def _f64(x):
    return np.ascontiguousarray(np.atleast_1d(np.asarray(x, dtype=np.float64)))


def _i32(x):
    return np.ascontiguousarray(np.atleast_1d(np.asarray(x, dtype=np.int32)))


# --------------------------------------------------------------- QNM table


def qnm_table():
    """PhenomD (2,2,0) QNM table and its scipy cubic-spline coefficients, as
    waveforms/phenomhm.py:148-163 prepares them (default not-a-knot spline)."""
    from scipy.interpolate import CubicSpline

    path = _HERE.parent / "bbhx_amd" / "waveforms" / "qnm_phenomd.txt"
    tab = np.loadtxt(path)
    a, fring, fdamp = tab[:, 0].copy(), tab[:, 1].copy(), tab[:, 2].copy()
    sr = CubicSpline(a, fring)
    sd = CubicSpline(a, fdamp)
    return dict(
        dspin=0.002,
        y_rd=fring.copy(), c1_rd=sr.c[-2].copy(), c2_rd=sr.c[-3].copy(), c3_rd=sr.c[-4].copy(),
        y_dm=fdamp.copy(), c1_dm=sd.c[-2].copy(), c2_dm=sd.c[-3].copy(), c3_dm=sd.c[-4].copy(),
    )


# --------------------------------------------------------------- kernels


def _phenomd_rd_args(tab):
    # the last segment of the table is padded so the 1001-entry coefficient
    # arrays can be indexed at the final node like the reference does
    out = []
    for key in ("y_rd", "c1_rd", "c2_rd", "c3_rd", "y_dm", "c1_dm", "c2_dm", "c3_dm"):
        v = np.zeros(1001)
        src = np.asarray(tab[key], dtype=np.float64)
        v[: src.size] = src
        out.append(v)
    return out


def get_phenomd_ringdown_frequencies(m1, m2, chi1z, chi2z, use_ref=False):
    """fRD, fDM of PhenomD in units of 1/M (PhenomHMWaveform.cu:1933-1993)."""
    m1, m2, chi1z, chi2z = map(_f64, (m1, m2, chi1z, chi2z))
    B = m1.size
    fr = np.zeros(B)
    fd = np.zeros(B)
    tab = qnm_table()
    arrs = _phenomd_rd_args(tab)
    lib = _load_ref() if use_ref else _load()
    fn = lib.ref_get_phenomd_ringdown_frequencies if use_ref else lib.orc_get_phenomd_ringdown_frequencies
    fn(_d(fr), _d(fd), _d(m1), _d(m2), _d(chi1z), _d(chi2z), C.c_int(B), *[_d(a) for a in arrs], C.c_double(tab["dspin"]))
    return fr, fd


def get_phenomhm_ringdown_frequencies(m1, m2, chi1z, chi2z, ells, mms, use_ref=False):
    """Per-mode QNM fRD, fDM in units of 1/M, layout (B, M) (PhenomHMWaveform.cu:1817-1877)."""
    m1, m2, chi1z, chi2z = map(_f64, (m1, m2, chi1z, chi2z))
    ells, mms = _i32(ells), _i32(mms)
    B, M = m1.size, ells.size
    fr = np.zeros(B * M)
    fd = np.zeros(B * M)
    lib = _load_ref() if use_ref else _load()
    fn = lib.ref_get_phenomhm_ringdown_frequencies if use_ref else lib.orc_get_phenomhm_ringdown_frequencies
    fn(_d(fr), _d(fd), _d(m1), _d(m2), _d(chi1z), _d(chi2z), _i(ells), _i(mms), C.c_int(M), C.c_int(B))
    return fr.reshape(B, M), fd.reshape(B, M)


def waveform_amp_phase(out, ells, mms, freqs, m1_SI, m2_SI, chi1z, chi2z, distance, f_ref,
                       num_modes, length, num_bin_all, Mf_RD, Mf_DM, run_phenomd, use_ref=False):
    """Fill amp | phase | tf slabs of ``out`` (PhenomHMWaveform.cu:2998)."""
    ells, mms = _i32(ells), _i32(mms)
    freqs = _f64(freqs)
    m1_SI, m2_SI, chi1z, chi2z, distance, f_ref = map(_f64, (m1_SI, m2_SI, chi1z, chi2z, distance, f_ref))
    Mf_RD, Mf_DM = _f64(Mf_RD), _f64(Mf_DM)
    assert out.dtype == np.float64 and out.flags.c_contiguous
    assert out.size >= 3 * num_bin_all * num_modes * length
    lib = _load_ref() if use_ref else _load()
    fn = lib.ref_waveform_amp_phase if use_ref else lib.orc_waveform_amp_phase
    fn(_d(out), _i(ells), _i(mms), _d(freqs), _d(m1_SI), _d(m2_SI), _d(chi1z), _d(chi2z), _d(distance),
       _d(f_ref), C.c_int(num_modes), C.c_int(length), C.c_int(num_bin_all), _d(Mf_RD), _d(Mf_DM),
       C.c_int(int(run_phenomd)))


def lisa_response(buffer, ells, mms, freqs, phi_ref, inc, lam, beta, psi, TDItag, rescaled, tdi2,
                  num_modes, length, num_bin_all, includes_amps, orbit):
    """Response.cu:671-746 with tabulated orbits.

    ``orbit`` is a dict with keys dt, N, n_arr (N,6,3), x_arr (N,3,3), links (6,), armlength.
    """
    ells, mms = _i32(ells), _i32(mms)
    freqs = _f64(freqs)
    phi_ref, inc, lam, beta, psi = map(_f64, (phi_ref, inc, lam, beta, psi))
    n_arr = np.ascontiguousarray(orbit["n_arr"], dtype=np.float64)
    x_arr = np.ascontiguousarray(orbit["x_arr"], dtype=np.float64)
    links = _i32(orbit["links"])
    assert buffer.dtype == np.float64 and buffer.flags.c_contiguous
    _load().orc_lisa_response(
        _d(buffer), _i(ells), _i(mms), _d(freqs), _d(phi_ref), _d(inc), _d(lam), _d(beta), _d(psi),
        C.c_int(TDItag), C.c_int(int(rescaled)), C.c_int(int(tdi2)), C.c_int(num_modes), C.c_int(length),
        C.c_int(num_bin_all), C.c_int(int(includes_amps)), C.c_double(orbit["dt"]), C.c_int(int(orbit["N"])),
        _d(n_arr), _d(x_arr), _i(links), C.c_double(orbit["armlength"]))


def sylm(l, m, theta, phi):
    """Spin-weight -2 spherical harmonic used by the response (Response.cu:383-424)."""
    out = np.zeros(2)
    _load().orc_sylm(C.c_int(l), C.c_int(m), C.c_double(theta), C.c_double(phi), _d(out))
    return complex(out[0], out[1])


def polarization(lam, beta, psi):
    """Wave vector k (3,) and polarisation tensors H+, Hx (3, 3) in the SSB frame
    (Response.cu:503-568)."""
    k, hp, hc = np.zeros(3), np.zeros(9), np.zeros(9)
    _load().orc_polarization(C.c_double(lam), C.c_double(beta), C.c_double(psi), _d(k), _d(hp), _d(hc))
    return k, hp.reshape(3, 3), hc.reshape(3, 3)


def interpolate(x, y, length, num_interp_params, num_modes, num_bin_all):
    """Not-a-knot spline coefficients (Interpolate.cu:320-359). Returns (B, c1, c2, c3)."""
    x, y = _f64(x), _f64(y)
    n = num_interp_params * num_modes * num_bin_all * length
    assert y.size == n and x.size == num_modes * num_bin_all * length
    B = np.zeros(n)
    c1 = np.zeros(n)
    c2 = np.zeros(n)
    c3 = np.zeros(n)
    _load().orc_interpolate(_d(x), _d(y), _d(B), _d(c1), _d(c2), _d(c3), C.c_int(length),
                            C.c_int(num_interp_params), C.c_int(num_modes), C.c_int(num_bin_all))
    return B, c1, c2, c3


def direct_sum(buffer, num_bin_all, length, num_modes, t_start, t_end, use_ref=False):
    """fill_waveform (WaveformBuild.cu:194-296). Returns (B, 3, L) complex."""
    buffer = _f64(buffer)
    t_start, t_end = _f64(t_start), _f64(t_end)
    out = np.zeros(num_bin_all * 3 * length, dtype=np.complex128)
    lib = _load_ref() if use_ref else _load()
    fn = lib.ref_direct_sum if use_ref else lib.orc_direct_sum
    fn(_d(out), _d(buffer), C.c_int(num_bin_all), C.c_int(length), C.c_int(3), C.c_int(num_modes),
       _d(t_start), _d(t_end))
    return out.reshape(num_bin_all, 3, length)


def interp_tdi(data_freqs, freqs, y, c1, c2, c3, t_start, t_end, length, num_bin_all, num_modes,
               inds_list, inds_start, ind_lengths, use_ref=False):
    """InterpTDI (WaveformBuild.cu:68-189). ``inds_list`` holds one int32 (M*n_b,)
    array per binary. Returns a list of (3, n_b) complex templates."""
    data_freqs, freqs, y, c1, c2, c3 = map(_f64, (data_freqs, freqs, y, c1, c2, c3))
    t_start, t_end = _f64(t_start), _f64(t_end)
    inds_start, ind_lengths = _i32(inds_start), _i32(ind_lengths)
    if use_ref:
        lib = _load_ref()
        inds_list = [np.ascontiguousarray(a, dtype=np.int32) for a in inds_list]
        templates = [np.zeros(3 * int(n), dtype=np.complex128) for n in ind_lengths]
        tptr = np.asarray([t.ctypes.data for t in templates], dtype=np.int64)
        iptr = np.asarray([a.ctypes.data for a in inds_list], dtype=np.int64)
        lib.ref_InterpTDI(_l(tptr), _d(data_freqs), _d(freqs), _d(y), _d(c1), _d(c2), _d(c3), _d(t_start),
                          _d(t_end), C.c_int(length), C.c_int(data_freqs.size), C.c_int(num_bin_all),
                          C.c_int(num_modes), _l(iptr), _i(inds_start), _i(ind_lengths))
        return [t.reshape(3, -1) for t in templates]
    toff = np.zeros(num_bin_all, dtype=np.int64)
    ioff = np.zeros(num_bin_all, dtype=np.int64)
    toff[1:] = np.cumsum(3 * ind_lengths.astype(np.int64))[:-1]
    ioff[1:] = np.cumsum(num_modes * ind_lengths.astype(np.int64))[:-1]
    flat = np.zeros(int(3 * ind_lengths.astype(np.int64).sum()), dtype=np.complex128)
    inds = _i32(np.concatenate(inds_list) if len(inds_list) else np.zeros(0, np.int32))
    _load().orc_interp_tdi(_d(flat), _l(toff), _d(data_freqs), _d(freqs), _d(y), _d(c1), _d(c2), _d(c3),
                           _d(t_start), _d(t_end), C.c_int(length), C.c_int(num_bin_all), C.c_int(num_modes),
                           _i(inds), _l(ioff), _i(inds_start), _i(ind_lengths))
    return [flat[o:o + 3 * int(n)].reshape(3, -1) for o, n in zip(toff, ind_lengths)]


def hdyn(residuals, data_constants, freqs, num_bin_all, length, include_last_segment=False, use_ref=False):
    """Heterodyned sums (Likelihood.cu:21-201). ``residuals`` flat (K, 3, B) complex,
    ``data_constants`` flat (4, 3, K) complex. Returns (d_h, h_h) complex (B,).

    ``use_ref`` runs the reference CPU loop, which always drops the last segment."""
    r = np.ascontiguousarray(residuals, dtype=np.complex128)
    dc = np.ascontiguousarray(data_constants, dtype=np.complex128)
    freqs = _f64(freqs)
    o1 = np.zeros(num_bin_all, dtype=np.complex128)
    o2 = np.zeros(num_bin_all, dtype=np.complex128)
    if use_ref:
        _load_ref().ref_hdyn(_d(o1), _d(o2), _d(r), _d(dc), _d(freqs), C.c_int(num_bin_all), C.c_int(length),
                             C.c_int(3))
    else:
        _load().orc_hdyn(_d(o1), _d(o2), _d(r), _d(dc), _d(freqs), C.c_int(num_bin_all), C.c_int(length),
                         C.c_int(3), C.c_int(int(include_last_segment)))
    return o1, o2


def direct_like(data_channels, noise_factors, templates, inds_start, ind_lengths, data_length, use_ref=False):
    """Direct inner products (Likelihood.cu:357-375, :478-529). ``templates`` is a list
    of (3, n_b) complex arrays, multiplied by the noise factors in place.
    Returns (d_h, h_h) complex (B,)."""
    d = np.ascontiguousarray(data_channels, dtype=np.complex128).ravel()
    w = _f64(noise_factors)
    inds_start, ind_lengths = _i32(inds_start), _i32(ind_lengths)
    B = len(templates)
    d_h = np.zeros(B, dtype=np.complex128)
    h_h = np.zeros(B, dtype=np.complex128)
    if use_ref:
        flats = [np.ascontiguousarray(t).ravel() for t in templates]
        ptrs = np.asarray([t.ctypes.data for t in flats], dtype=np.int64)
        _load_ref().ref_direct_like(_d(d_h), _d(h_h), _d(d), _d(w), _l(ptrs), _i(inds_start), _i(ind_lengths),
                                    C.c_int(data_length), C.c_int(B))
        for t, f in zip(templates, flats):
            t[...] = f.reshape(t.shape)
        return d_h, h_h
    toff = np.zeros(B, dtype=np.int64)
    toff[1:] = np.cumsum(3 * ind_lengths.astype(np.int64))[:-1]
    flat = np.concatenate([np.ascontiguousarray(t).ravel() for t in templates]) if B else np.zeros(0, np.complex128)
    flat = np.ascontiguousarray(flat, dtype=np.complex128)
    _load().orc_direct_like(_d(d_h), _d(h_h), _d(d), _d(w), _d(flat), _l(toff), _i(inds_start), _i(ind_lengths),
                            C.c_int(data_length), C.c_int(B))
    for t, o in zip(templates, toff):
        t[...] = flat[o:o + t.size].reshape(t.shape)
    return d_h, h_h
