"""numpy restatement of the reference's host-side logic around the kernels.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows, step by step and in numpy:
  waveforms/phenomhm.py:270-614   (PhenomHMAmpPhase.run_wave / __call__)
  response/fastfdresponse.py:278-531 (LISATDIResponse.__call__)
  waveformbuild.py:89-239, :338-700  (TemplateInterpFD, BBHWaveformFD)
  likelihood.py:26-225, :380-648     (Likelihood, HeterodynedLikelihood)
  utils/transform.py:75-105          (tSSBfromLframe)
with the oracle kernels of this package in place of the reference's backend.
"""

from __future__ import annotations

import numpy as np

from . import (direct_like, direct_sum, get_phenomd_ringdown_frequencies, get_phenomhm_ringdown_frequencies, hdyn,
               interp_tdi, interpolate, lisa_response, waveform_amp_phase)

MSUN_SI = 1.98848e+30
YRSID_SI = 31558149.763545603
AU_SI = 149597870700.0
C_SI = 299792458.
MTSUN_SI = 4.925491025873693e-06
ConstOmega = 1.99098659277e-7

ALL_MODES = [(2, 2), (3, 3), (4, 4), (2, 1), (3, 2), (4, 3)]


def tSSBfromLframe(tL, lambdaSSB, betaSSB, t0=0.0):
    """utils/transform.py:75-105."""
    ConstPhi0 = ConstOmega * t0
    phase = ConstOmega * tL + ConstPhi0 - lambdaSSB
    RoC = AU_SI / C_SI
    return tL + RoC * np.cos(betaSSB) * np.cos(phase) - 1.0 / 2 * ConstOmega * pow(RoC * np.cos(betaSSB), 2) * np.sin(
        2.0 * phase)


class OracleWaveform:
    """BBHWaveformFD with PhenomHMAmpPhase and LISATDIResponse, in numpy."""

    def __init__(self, orbit, run_phenomd=False, mf_min=1e-4, mf_max=0.6, initial_t_val=0.0, TDItag="AET",
                 rescaled=False, tdi2=False):
        self.orbit = orbit
        self.run_phenomd = run_phenomd
        self.mf_min, self.mf_max, self.initial_t_val = mf_min, mf_max, initial_t_val
        self.TDItag_int = 1 if TDItag == "XYZ" else 2
        self.rescaled, self.tdi2 = rescaled, tdi2
        self.t_base_max = (orbit["N"] - 1) * orbit["dt"]

    # ---- phenomhm.py:270-507
    def run_wave(self, m1, m2, chi1z, chi2z, distance, phi_ref, f_ref, t_ref, length, freqs=None, buffer=None,
                 modes=None):
        m1, m2, chi1z, chi2z = (np.array(np.atleast_1d(v), dtype=np.float64) for v in (m1, m2, chi1z, chi2z))
        distance, f_ref, t_ref = (np.array(np.atleast_1d(v), dtype=np.float64) for v in (distance, f_ref, t_ref))
        if np.any(np.abs(chi1z) > 1.0) or np.any(np.abs(chi2z) > 1.0):
            raise ValueError("spin magnitude > 1")
        sw = m1 < m2
        m1[sw], m2[sw] = m2[sw].copy(), m1[sw].copy()
        chi1z[sw], chi2z[sw] = chi2z[sw].copy(), chi1z[sw].copy()
        modes = ALL_MODES if modes is None else modes
        if self.run_phenomd:
            modes = [(2, 2)]
        ells = np.asarray([l for l, m in modes], np.int32)
        mms = np.asarray([m for l, m in modes], np.int32)
        B, M, L = len(m1), len(ells), int(length)
        self.B, self.M, self.L, self.modes = B, M, L, modes
        nper = B * M * L
        if buffer is None:
            buffer = np.zeros(3 * nper)
        if freqs is None:
            Msec = (m1 + m2) * MTSUN_SI
            base = np.logspace(np.log10(self.mf_min), np.log10(self.mf_max), L)
            freqs = (base[None, None, :] / Msec[:, None, None] * mms[None, :, None] / 2).flatten()
        elif freqs.ndim == 1:
            freqs = np.tile(freqs, (B, M, 1)).flatten()
        else:
            freqs = freqs.flatten().copy()
        self.freqs = freqs
        frd, fdm = get_phenomd_ringdown_frequencies(m1, m2, chi1z, chi2z)
        if self.run_phenomd:
            Mf_RD, Mf_DM = frd, fdm
        else:
            hr, hd = get_phenomhm_ringdown_frequencies(m1, m2, chi1z, chi2z, ells, mms)
            Mf_RD = np.concatenate([hr, frd[:, None]], axis=1).flatten()
            Mf_DM = np.concatenate([hd, fdm[:, None]], axis=1).flatten()
        waveform_amp_phase(buffer, ells, mms, freqs, m1 * MSUN_SI, m2 * MSUN_SI, chi1z, chi2z, distance, f_ref, M, L,
                           B, Mf_RD, Mf_DM, self.run_phenomd)
        temp = freqs.reshape(B, M, -1) * (t_ref[:, None, None] - self.initial_t_val) * 2 * np.pi
        buffer[nper:2 * nper] += temp.flatten()
        buffer[2 * nper:3 * nper] += np.tile(t_ref[:, None, None], (1, M, L)).flatten()
        self.buffer = buffer
        return buffer

    def amp_phase(self, *args, Tobs=None, direct=False, **kwargs):
        """phenomhm.py:509-614."""
        self.run_wave(*args, **kwargs)
        if direct:
            return
        B, M, L = self.B, self.M, self.L
        nper = B * M * L
        tf = self.buffer[2 * nper:3 * nper].reshape(B, M, L)
        fs = self.freqs.reshape(B, M, L)
        fix = tf.min(axis=-1) < 0.0
        f_min = fs.min(axis=-1)
        if np.any(fix):
            w = np.argwhere((0.0 < tf[fix, 1:]) & (0.0 > tf[fix, :-1]))
            if w.shape[0] != fix.sum():
                raise ValueError("window search at x_min failed")
            s = w[:, 1]
            y1 = np.take_along_axis(fs[fix], s[:, None], axis=-1)[:, 0]
            y2 = np.take_along_axis(fs[fix], s[:, None] + 1, axis=-1)[:, 0]
            x1 = np.take_along_axis(tf[fix], s[:, None], axis=-1)[:, 0]
            x2 = np.take_along_axis(tf[fix], s[:, None] + 1, axis=-1)[:, 0]
            f_min[fix] = (y2 - y1) / (x2 - x1) * (0.0 - x1) + y1
        if Tobs is None:
            Tobs = np.full((B, M), 1e300)
        elif np.ndim(Tobs) == 1:
            Tobs = np.repeat(np.asarray(Tobs)[:, None], M, axis=-1)
        fix = tf.max(axis=-1) > Tobs
        f_max = fs.max(axis=-1)
        if np.any(fix):
            w = np.argwhere((Tobs[fix][:, None] < tf[fix, 1:]) & (Tobs[fix][:, None] > tf[fix, :-1]))
            if w.shape[0] != fix.sum():
                raise ValueError("window search at x_max failed")
            s = w[:, 1]
            y1 = np.take_along_axis(fs[fix], s[:, None], axis=-1)[:, 0]
            y2 = np.take_along_axis(fs[fix], s[:, None] + 1, axis=-1)[:, 0]
            x1 = np.take_along_axis(tf[fix], s[:, None], axis=-1)[:, 0]
            x2 = np.take_along_axis(tf[fix], s[:, None] + 1, axis=-1)[:, 0]
            f_max[fix] = (y2 - y1) / (x2 - x1) * (Tobs[fix] - x1) + y1
        new_freqs = np.logspace(np.log10(f_min), np.log10(f_max), L, axis=-1)
        kwargs = dict(kwargs)
        kwargs["freqs"] = new_freqs
        self.run_wave(*args, **kwargs)

    # ---- waveformbuild.py:338-700
    def __call__(self, m1, m2, chi1z, chi2z, distance, phi_ref, f_ref, inc, lam, beta, psi, t_ref, t_obs_start=0.0,
                 t_obs_end=1.0, freqs=None, length=None, modes=None, direct=False, compress=True, squeeze=False,
                 fill=False, combine=False):
        lam, beta = np.atleast_1d(lam).astype(float), np.atleast_1d(beta).astype(float)
        m1 = np.atleast_1d(m1).astype(float)
        B = len(m1)
        t_start = np.atleast_1d(tSSBfromLframe(t_obs_start * YRSID_SI, lam, beta, 0.0))
        t_end = np.atleast_1d(tSSBfromLframe(t_obs_end * YRSID_SI, lam, beta, 0.0))
        if direct:
            length = len(freqs)
        modes = ALL_MODES if modes is None else modes
        M = 1 if self.run_phenomd else len(modes)
        buf = np.zeros(9 * length * M * B)
        self.amp_phase(m1, m2, chi1z, chi2z, distance, np.zeros(B), f_ref, t_ref, length,
                       freqs=freqs if direct else None, buffer=buf, modes=modes, Tobs=t_end - t_start, direct=direct)
        nper = B * M * length
        if not direct:
            tf = buf[2 * nper:3 * nper]
            if tf.min() < 0.0 or tf.max() > self.t_base_max:
                raise ValueError("orbit coverage")
        ells = np.asarray([l for l, m in self.modes], np.int32)
        mms = np.asarray([m for l, m in self.modes], np.int32)
        lisa_response(buf, ells, mms, self.freqs, np.broadcast_to(np.atleast_1d(phi_ref), (B,)).copy(),
                      np.broadcast_to(np.atleast_1d(inc), (B,)).copy(), lam, beta,
                      np.broadcast_to(np.atleast_1d(psi), (B,)).copy(), self.TDItag_int, self.rescaled, self.tdi2,
                      M, length, B, 1, self.orbit)
        self.buffer_final = buf.reshape(9, B, M, length).copy()
        self.t_start, self.t_end = t_start, t_end
        if direct and compress:
            out = direct_sum(buf, B, length, M, t_start, t_end)
            return out.squeeze() if squeeze else out
        if direct:
            # per-mode direct outputs (waveformbuild.py:603-636): the direct sum
            # of each mode's slice of the buffer on its own
            bf = buf.reshape(9, B, M, length)
            out = np.zeros((B, 3, M, length), dtype=np.complex128)
            for mi in range(M):
                out[:, :, mi, :] = direct_sum(np.ascontiguousarray(bf[:, :, mi, :]).ravel(), B, length, 1, t_start,
                                              t_end)
            return out.squeeze() if squeeze else out
        # spline + dense interpolation (waveformbuild.py:638-700)
        _, c1, c2, c3 = interpolate(self.freqs, buf, length, 9, M, B)
        data_freqs = np.asarray(freqs, dtype=np.float64)
        fsh = self.freqs.reshape(B, M, length)
        st = np.searchsorted(data_freqs, fsh.min(axis=(1, 2)), side="right")
        et = np.searchsorted(data_freqs, fsh.max(axis=(1, 2)), side="right")
        inds = [np.concatenate([np.searchsorted(fsh[i, j], data_freqs[st[i]:et[i]], side="right") - 1
                                for j in range(M)]).astype(np.int32) for i in range(B)]
        lengths = (et - st).astype(np.int32)
        tmpl = interp_tdi(data_freqs, self.freqs, buf, c1, c2, c3, t_start, t_end, length, B, M, inds,
                          st.astype(np.int32), lengths)
        self.start_inds, self.lengths = st.astype(np.int32), lengths
        if fill:
            if combine:
                out = np.zeros((3, len(data_freqs)), dtype=np.complex128)
                for t, s, n in zip(tmpl, st, lengths):
                    out[:, s:s + n] += t
                return out.squeeze() if squeeze else out
            out = np.zeros((B, 3, len(data_freqs)), dtype=np.complex128)
            for b, (t, s, n) in enumerate(zip(tmpl, st, lengths)):
                out[b, :, s:s + n] = t
            return out
        return tmpl, self.start_inds, self.lengths


class OracleLikelihood:
    """likelihood.py:26-225."""

    def __init__(self, wave, data_freqs, data_channels, psd):
        self.wave = wave
        self.data_freqs = np.asarray(data_freqs, dtype=np.float64)
        psd = np.asarray(psd, dtype=np.float64)
        df = np.zeros_like(psd)
        df[:, 1:] = np.diff(self.data_freqs)
        df[:, 0] = df[:, 1]
        self.noise_factors = np.sqrt(1.0 / psd * df).flatten()
        self.data_channels = self.noise_factors * np.asarray(data_channels).flatten()
        self.d_d = (4 * np.sum(self.data_channels.conj() * self.data_channels).real).item()

    def get_ll(self, params, **kw):
        kw = dict(kw)
        kw.update(freqs=self.data_freqs, fill=False, direct=False)
        tmpl, st, n = self.wave(*params, **kw)
        d_h, h_h = direct_like(self.data_channels, self.noise_factors, tmpl, st, n, len(self.data_freqs))
        return -1 / 2 * (self.d_d + h_h - 2 * d_h).real, d_h, h_h


def hdyn_constants(f_dense, d, h0, S_n, freqs):
    """Heterodyne constants of a reference template (likelihood.py:499-527).

    f_dense (N,) dense frequencies already cut to [freqs[0], freqs[-1]]; d, h0
    (3, N) complex data and reference template there; S_n (3, N) PSD; freqs (K,)
    sparse nodes. Returns the flat (4, 3, K) complex constants A0, A1, B0, B1
    (segment j = [freqs[j], freqs[j+1]) summed into slot j + 1, the last dense
    point left out, as the reference's loop over unique(bins[:-1]) does) and the
    totals (d_d, h_h, Re d_h) over the whole band (likelihood.py:529-533)."""
    K = freqs.size
    bins = np.searchsorted(freqs, f_dense, "right") - 1
    f_m = (freqs[1:] + freqs[:-1]) / 2
    df = f_dense[1] - f_dense[0]
    A0 = 4 * (h0.conj() * d) / S_n * df
    A1 = A0 * (f_dense - f_m[bins])
    B0 = 4 * (h0.conj() * h0) / S_n * df
    B1 = B0 * (f_dense - f_m[bins])
    C = [np.zeros((3, K), dtype=np.complex128) for _ in range(4)]
    for ind in np.unique(bins[:-1]):
        keep = bins == ind
        keep[-1] = False
        for c, flat in zip(C, (A0, A1, B0, B1)):
            c[:, ind + 1] = np.sum(flat[:, keep], axis=1)
    totals = (np.sum(4 * (d.conj() * d) / S_n * df).real, np.sum(B0).real, np.sum(A0).real)
    return np.concatenate([c.flatten() for c in C]), totals


class OracleHeterodyne:
    """likelihood.py:380-648, with the constants built as the reference's Python loop does."""

    def __init__(self, wave, data_freqs, data_channels, ref_params, K, psd_fn, reference_length=8096,
                 include_last_segment=True, **wkw):
        self.wave = wave
        self.include_last = include_last_segment
        f_dense = np.asarray(data_freqs, dtype=np.float64)
        d = np.asarray(data_channels)
        minF = f_dense.min() * 0.999999999999
        maxF = f_dense.max() * 1.000000000001
        if minF == 0.0:
            minF = 1e-6
        freqs = np.logspace(np.log10(minF), np.log10(maxF), K)
        h0 = wave(*ref_params, freqs=f_dense, length=reference_length, direct=False, fill=True, compress=True,
                  squeeze=True, **wkw)[0]
        h0_temp = wave(*ref_params, freqs=freqs, direct=True, compress=True, squeeze=True, **wkw)[0]
        fk = freqs[~(np.abs(h0_temp) == 0.0)]
        freqs = np.logspace(np.log10(fk[0]), np.log10(fk[-1]), K)
        self.h0_sparse = wave(*ref_params, freqs=freqs, direct=True, compress=True, squeeze=True, **wkw)[None, :, :]
        inds = (f_dense >= freqs[0]) & (f_dense <= freqs[-1])
        f_dense, d, h0 = f_dense[inds], d[:, inds], h0[:, inds]
        S_n = np.asarray(psd_fn(f_dense))
        self.data_constants, (self.reference_d_d, self.reference_h_h, self.reference_d_h) = hdyn_constants(
            f_dense, d, h0, S_n, freqs)
        self.reference_ll = -0.5 * (self.reference_d_d + self.reference_h_h - 2 * self.reference_d_h)
        self.freqs = freqs
        self.wkw = wkw

    def get_ll(self, params):
        h = self.wave(*params, freqs=self.freqs, direct=True, compress=True, squeeze=False, **self.wkw)
        r = h / self.h0_sparse
        B = r.shape[0]
        res = r.transpose((2, 1, 0)).flatten()
        d_h, h_h = hdyn(res, self.data_constants, self.freqs, B, len(self.freqs),
                        include_last_segment=self.include_last)
        return -0.5 * (self.reference_d_d + h_h - 2 * d_h).real, d_h, h_h
