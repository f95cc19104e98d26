"""Direct and heterodyned log-likelihoods on the GPU.

Host-side mirror of the reference's ``likelihood.py`` (``Likelihood`` :26-225,
``HeterodynedLikelihood`` :228-648): same constructor and ``get_ll``
signatures and semantics. ``HeterodynedLikelihood.get_ll`` runs the fused
single-pass kernel by default (``fused=True``); ``fused=False`` runs the
kernel-by-kernel path of the reference (waveform -> response -> direct sum ->
r = h/h0 -> hdyn).
"""

from __future__ import annotations

import numpy as np
import torch

from .utils.arrays import as_np1d, to_dev, to_host
from .utils.parallelbase import BBHxParallelModule
from .utils.constants import YRSID_SI
from .utils.sensitivity import AET1SensitivityMatrix


class Likelihood(BBHxParallelModule):
    """log L = -1/2 (<d|d> + <h|h> - 2 <d|h>) summed over each template's band
    (likelihood.py:26-225). ``data_channels`` and ``psd`` are (3, N)."""

    def __init__(self, template_gen, data_freqs, data_channels, psd, force_backend=None):
        super().__init__(force_backend=force_backend)
        torch = self.xp
        data_freqs_cpu = np.asarray(to_host(data_freqs), dtype=np.float64)
        self.data_freqs = to_dev(data_freqs_cpu, self.device)
        psd = np.asarray(to_host(psd), dtype=np.float64)
        self.psd = psd
        self.delta_f = np.zeros_like(self.psd)
        self.delta_f[:, 1:] = np.diff(data_freqs_cpu)
        self.delta_f[:, 0] = self.delta_f[:, 1]
        self.noise_factors = to_dev(np.sqrt(1.0 / psd * self.delta_f).flatten(), self.device)
        dc = to_dev(data_channels, self.device, dtype=torch.complex128).flatten()
        self.data_channels = (self.noise_factors * dc).contiguous()
        self.waveform_gen = template_gen
        self.data_stream_length = len(data_freqs_cpu)
        self.d_d = (4 * torch.sum(self.data_channels.conj() * self.data_channels).real).item()

    @property
    def like_gen(self):
        return self.backend.direct_like_wrap

    def get_ll(self, params, return_extracted_snr=False, phase_marginalize=False, **waveform_kwargs):
        """params: (12,) or (12, num_bin_all) rows m1, m2, chi1z, chi2z, distance,
        phi_ref, f_ref, inc, lam, beta, psi, t_ref (likelihood.py:140-225)."""
        torch = self.xp
        self.phase_marginalize = phase_marginalize
        self.return_extracted_snr = return_extracted_snr
        waveform_kwargs["freqs"] = self.data_freqs
        waveform_kwargs["fill"] = False
        waveform_kwargs["direct"] = False
        templateChannels, inds_start, ind_lengths = self.waveform_gen(*params, **waveform_kwargs)
        templateChannels = [tc.flatten() for tc in templateChannels]
        ptrs = np.asarray([tc.data_ptr() for tc in templateChannels], dtype=np.int64)
        B = self.waveform_gen.num_bin_all
        self.d_h = torch.zeros(B, dtype=torch.complex128, device=self.device)
        self.h_h = torch.zeros(B, dtype=torch.complex128, device=self.device)
        self.like_gen(self.d_h, self.h_h, self.data_channels, self.noise_factors, ptrs, inds_start, ind_lengths,
                      self.data_stream_length, B)
        d_h_temp = self.d_h if not self.phase_marginalize else torch.abs(self.d_h)
        out = (-1 / 2 * (self.d_d + self.h_h - 2 * d_h_temp).real).cpu().numpy()
        if self.return_extracted_snr:
            snr = (d_h_temp.real / torch.sqrt(self.h_h.real)).cpu().numpy()
            return np.array([out, snr]).T
        return out


    # ---------------------------------------------------------- fused ragged path

    def direct_plan(self, lengths, modes=None, chunk=16384, block=256, grid=None, min_chunk=1024):
        """Plan (sizes + device buffers) of the fused direct likelihood for a
        batch whose binary b has a sparse grid of lengths[b] points."""
        gen = self.waveform_gen
        M = len(modes) if modes is not None else len(gen.amp_phase_gen.allowable_modes)
        return self.backend.DirectPlan(lengths, M, self.data_stream_length, chunk=chunk, block=block, grid=grid,
                                       min_chunk=min_chunk)

    def get_ll_fused_device(self, plan, rows, t_start, t_end, flags, modes=None, precision=None):
        """Fused direct log-likelihood of device inputs (``fused_inputs``): a
        device tensor, with no host synchronisation (graph-capturable; flags
        must be zeroed before the call). precision (None: the plan's current
        one): 0 FP64, 1 FP32 bulk with FP64 phase (``DirectPlan.set_precision``;
        the first switch to 1 allocates, so make it before a graph capture)."""
        torch = self.xp
        if precision is not None:
            plan.set_precision(precision)
        gen = self.waveform_gen
        if modes is None:
            modes = gen.amp_phase_gen.allowable_modes
        ells = np.asarray([l for l, m in modes], dtype=np.int32)
        mms = np.asarray([m for l, m in modes], dtype=np.int32)
        B = plan.num_bin_all
        if not hasattr(self, "_dfused_out") or self._dfused_out[0].numel() != B:
            self._dfused_out = (torch.empty(B, dtype=torch.complex128, device=self.device),
                                torch.empty(B, dtype=torch.complex128, device=self.device))
        d_h, h_h = self._dfused_out
        ap, resp = gen.amp_phase_gen, gen.response_gen
        self.backend.direct_fused(plan, d_h, h_h, flags, rows, t_start, t_end, ells, mms, self.data_freqs,
                                  self.data_channels, self.noise_factors, ap.qnm_packed, ap.dspin, ap.initial_t_val,
                                  ap.mf_min, ap.mf_max, resp.cpp_response, resp.TDItag_int, resp.rescaled, resp.tdi2)
        self.d_h, self.h_h = d_h, h_h
        return -1 / 2 * (self.d_d + h_h - 2 * d_h).real

    def get_ll_fused(self, params, lengths, return_extracted_snr=False, phase_marginalize=False, plan=None,
                     modes=None, t_obs_start=0.0, t_obs_end=1.0, precision=0):
        """Direct log-likelihood of templates with per-template sparse grid
        lengths (``lengths[b]``), fused: no template is materialised. Same
        params layout and result as ``get_ll(params, length=...)``.
        precision 1: the FP32-bulk form (``get_ll_fused_device``)."""
        torch = self.xp
        rows, t_start, t_end = fused_param_rows(self.waveform_gen, params, t_obs_start, t_obs_end, self.device)
        lengths = np.asarray(lengths, dtype=np.int32)
        if lengths.shape != (rows.shape[1],):
            raise ValueError("lengths must have one entry per template")
        if plan is None:
            plan = self.direct_plan(lengths, modes)
        flags = torch.zeros(rows.shape[1], dtype=torch.int32, device=self.device)
        ll = self.get_ll_fused_device(plan, rows, t_start, t_end, flags, modes, precision=precision)
        fl = flags.cpu().numpy()
        if np.any(fl & 1):
            raise ValueError("The search for a window at x_min did not run correctly. If t_ref is close to zero, "
                             "the time-frequency correspondence may be crossing zero twice.")
        if np.any(fl & 2):
            raise ValueError("The search for a window at x_max did not run correctly. If t_ref is close to the "
                             "observation time, the time-frequency correspondence may be crossing zero twice.")
        d_h_temp = self.d_h if not phase_marginalize else torch.abs(self.d_h)
        out = (-1 / 2 * (self.d_d + self.h_h - 2 * d_h_temp).real).cpu().numpy()
        if return_extracted_snr:
            snr = (d_h_temp.real / torch.sqrt(self.h_h.real)).cpu().numpy()
            return np.array([out, snr]).T
        return out


    # ---------------------------------------------------------- gradient (fused ragged path)

    GRAD_COORDS = ("lnM", "q", "chi1z", "chi2z", "distance", "phi_ref", "cos_inc", "lam", "sin_beta", "psi", "t_ref")

    def get_ll_grad_fused_device(self, plan, rows, t_start, t_end, flags, modes=None):
        """Fused direct log-likelihood (B,) and its gradient (11, B) in the
        sampling coordinates (``GRAD_COORDS``), from device inputs; no host
        synchronisation. The plan's gradient buffers are allocated on first use.
        The observation window and the sparse grids are held fixed."""
        torch = self.xp
        gen = self.waveform_gen
        if modes is None:
            modes = gen.amp_phase_gen.allowable_modes
        ells = np.asarray([l for l, m in modes], dtype=np.int32)
        mms = np.asarray([m for l, m in modes], dtype=np.int32)
        B = plan.num_bin_all
        if not getattr(plan, "_grad_ready", False):
            plan.enable_grad()
            plan._grad_ready = True
        d_h = torch.empty(B, dtype=torch.complex128, device=self.device)
        h_h = torch.empty(B, dtype=torch.complex128, device=self.device)
        grad = torch.empty((11, B), dtype=torch.float64, device=self.device)
        ap, resp = gen.amp_phase_gen, gen.response_gen
        self.backend.direct_fused_grad(plan, d_h, h_h, grad, flags, rows, t_start, t_end, ells, mms, self.data_freqs,
                                       self.data_channels, self.noise_factors, ap.qnm_packed, ap.dspin,
                                       ap.initial_t_val, ap.mf_min, ap.mf_max, resp.cpp_response, resp.TDItag_int,
                                       resp.rescaled, resp.tdi2)
        self.d_h, self.h_h = d_h, h_h
        return -1 / 2 * (self.d_d + h_h - 2 * d_h).real, grad

    def get_ll_and_grad_fused(self, x, lengths, plan=None, modes=None, t_obs_start=0.0, t_obs_end=1.0):
        """Direct log-likelihood (B,) and gradient (B, 11) at sampling
        coordinates x (B, 11) (see ``HeterodynedLikelihood.sampling_to_params``)
        for templates with sparse grid lengths ``lengths``; numpy arrays."""
        torch = self.xp
        params = HeterodynedLikelihood.sampling_to_params(np.asarray(x, dtype=np.float64))
        rows, t_start, t_end = fused_param_rows(self.waveform_gen, params, t_obs_start, t_obs_end, self.device)
        lengths = np.asarray(lengths, dtype=np.int32)
        if lengths.shape != (rows.shape[1],):
            raise ValueError("lengths must have one entry per template")
        if plan is None:
            plan = self.direct_plan(lengths, modes)
        flags = torch.zeros(rows.shape[1], dtype=torch.int32, device=self.device)
        ll, grad = self.get_ll_grad_fused_device(plan, rows, t_start, t_end, flags, modes)
        if np.any(flags.cpu().numpy() != 0):
            raise ValueError("The search for an observation window did not run correctly for some template.")
        return ll.cpu().numpy(), grad.T.cpu().numpy()

    def loglike(self, x, lengths, plan=None, modes=None):
        """Differentiable direct log-likelihood: x (B, 11) float64 torch tensor of
        sampling coordinates; backward through the adjoint kernels (``DirectLogLike``)."""
        return DirectLogLike.apply(x, self, lengths, plan, modes)


def fused_param_rows(template_gen, params, t_obs_start, t_obs_end, device):
    """Device inputs of the fused kernels: (11, B) parameter rows (m1 >= m2
    enforced as phenomhm.py:175-199 does; f_ref dropped), t_start, t_end
    (SSB-frame window, waveformbuild.py:477-484)."""
    p = np.atleast_2d(np.asarray(to_host(params), dtype=np.float64))
    if p.shape[0] != 12:
        p = p.T
    if p.shape[0] != 12:
        raise ValueError("params must have 12 rows (m1, m2, chi1z, chi2z, distance, phi_ref, f_ref, inc, lam, "
                         "beta, psi, t_ref)")
    m1, m2, c1, c2 = p[0].copy(), p[1].copy(), p[2].copy(), p[3].copy()
    template_gen.amp_phase_gen._sanity_check_params(m1, m2, c1, c2)
    B = p.shape[1]
    rows = [m1, m2, c1, c2, p[4], p[5], p[7], p[8], p[9], p[10], p[11]]
    if np.isscalar(t_obs_start) and np.isscalar(t_obs_end):
        # one host-to-device copy of the 11 rows; the window from lam, beta on
        # the device (the reference evaluates it with its array module there)
        dev = to_dev(np.stack(rows), device)
        t_start = torch.empty(B, dtype=torch.float64, device=dev.device)
        t_end = torch.empty(B, dtype=torch.float64, device=dev.device)
        template_gen.backend.window_times(t_start, t_end, dev[7], dev[8], float(t_obs_start) * YRSID_SI,
                                          float(t_obs_end) * YRSID_SI)
        return dev, t_start, t_end
    # per-binary windows: evaluated with the rows on the host (time_window
    # checks their types), sent as one (13, B) block
    t_start, t_end = template_gen.time_window(p[8], p[9], t_obs_start, t_obs_end)
    block = np.stack(rows + [np.broadcast_to(t_start, (B,)), np.broadcast_to(t_end, (B,))])
    dev = to_dev(block, device)
    return dev[:11], dev[11], dev[12]


def sampling_rows(x):
    """(11, B) parameter rows of the fused kernels from sampling coordinates x
    (B, 11) (as ``sampling_rows_device``), with torch operations on x's device:
    m1 >= m2 enforced with the spins swapped along (phenomhm.py:175-199)."""
    M = torch.exp(x[:, 0])
    q = x[:, 1]
    m1, m2 = M * q / (1 + q), M / (1 + q)
    c1, c2 = x[:, 2], x[:, 3]
    swap = m1 < m2
    return torch.stack([torch.where(swap, m2, m1), torch.where(swap, m1, m2), torch.where(swap, c2, c1),
                        torch.where(swap, c1, c2), x[:, 4], x[:, 5], torch.arccos(x[:, 6]), x[:, 7],
                        torch.arcsin(x[:, 8]), x[:, 9], x[:, 10]]).contiguous()


def sampling_rows_device(x, t_obs_start=0.0, t_obs_end=1.0):
    """Device inputs of the fused kernels from sampling coordinates x (B, 11)
    (ln(m1 + m2), q = m1/m2, chi1z, chi2z, distance, phi_ref, cos inc, lam,
    sin beta, psi, t_ref; a float64 tensor on the GPU): (11, B) rows with
    m1 >= m2 enforced (phenomhm.py:175-199) and the SSB-frame window
    (waveformbuild.py:477-484, the library's window kernel), computed on the
    device only, so that nothing waits for the host and the call can sit
    inside a captured graph. Unlike ``fused_param_rows`` the spins are not checked on
    the host (|chi| <= 1 is the caller's to keep)."""
    from .cutils import window_times

    rows = sampling_rows(x)
    t_start = torch.empty(rows.shape[1], dtype=torch.float64, device=rows.device)
    t_end = torch.empty_like(t_start)
    window_times(t_start, t_end, rows[7], rows[8], float(t_obs_start) * YRSID_SI, float(t_obs_end) * YRSID_SI)
    return rows, t_start, t_end


class HeterodynedLikelihood(BBHxParallelModule):
    """Heterodyned (relative-binning) log-likelihood (likelihood.py:228-648;
    arXiv:1806.08792).

    Extra keyword (no reference counterpart):
        hdyn_last_bin (str): ``"gpu"`` (default) uses all K-1 sparse segments like
            the reference's GPU kernel; ``"cpu"`` drops the last one like the
            reference's CPU loop (Likelihood.cu:150).
    """

    def __init__(self, template_gen, data_freqs, data_channels, reference_template_params, length_f_het,
                 template_gen_kwargs=None, reference_gen_kwargs=None, sens_mat=None, force_backend=None,
                 hdyn_last_bin="gpu"):
        template_gen_kwargs = {} if template_gen_kwargs is None else template_gen_kwargs
        reference_gen_kwargs = {} if reference_gen_kwargs is None else reference_gen_kwargs
        super().__init__(force_backend=force_backend)
        self.template_gen = template_gen
        self.f_dense = to_dev(data_freqs, self.device)
        self.d = to_dev(data_channels, self.device, dtype=self.xp.complex128)
        self.length_f_het = length_f_het
        if hdyn_last_bin not in ("gpu", "cpu"):
            raise ValueError("hdyn_last_bin must be 'gpu' or 'cpu'")
        self.include_last_segment = hdyn_last_bin == "gpu"
        self.sens_mat = sens_mat
        self.init_heterodyne_info(reference_template_params, template_gen_kwargs=template_gen_kwargs,
                                  reference_gen_kwargs=reference_gen_kwargs)

    @property
    def sens_mat(self):
        return self._sens_mat

    @sens_mat.setter
    def sens_mat(self, sens_mat):
        if sens_mat is None:
            sens_mat = AET1SensitivityMatrix(np.logspace(-5, -1, 1000))
        self._sens_mat = sens_mat

    @property
    def like_gen(self):
        return self.backend.hdyn_wrap

    def init_heterodyne_info(self, reference_template_params, template_gen_kwargs=None, reference_gen_kwargs=None):
        """Reference template and heterodyne constants (likelihood.py:380-548)."""
        torch = self.xp
        template_gen_kwargs = {} if template_gen_kwargs is None else template_gen_kwargs
        reference_gen_kwargs = {} if reference_gen_kwargs is None else reference_gen_kwargs
        template_gen_kwargs["squeeze"] = True
        template_gen_kwargs["compress"] = True
        template_gen_kwargs["direct"] = True

        minF = float(self.f_dense.min()) * 0.999999999999
        maxF = float(self.f_dense.max()) * 1.000000000001
        if minF == 0.0:
            minF = 1e-6

        reference_gen_kwargs["squeeze"] = True
        reference_gen_kwargs["compress"] = True
        reference_gen_kwargs["fill"] = True
        if not ("direct" in reference_gen_kwargs and reference_gen_kwargs["direct"] is True):
            reference_gen_kwargs["direct"] = False
            if "length" not in reference_gen_kwargs:
                reference_gen_kwargs["length"] = 8096

        freqs = np.logspace(np.log10(minF), np.log10(maxF), self.length_f_het)
        self.reference_template_params = np.asarray(reference_template_params, dtype=np.float64)
        h0 = self.template_gen(*reference_template_params, freqs=self.f_dense, **reference_gen_kwargs)[0]
        h0_temp = self.template_gen(*reference_template_params, freqs=freqs, **template_gen_kwargs)[0]
        keep = ~(torch.abs(h0_temp) == 0.0).cpu().numpy()
        freqs_keep = freqs[keep]
        freqs = np.logspace(np.log10(freqs_keep[0]), np.log10(freqs_keep[-1]), self.length_f_het)
        self.h0_sparse = self.template_gen(*reference_template_params, freqs=freqs, **template_gen_kwargs)[None, :, :]
        # the fused kernels read h0 as (3, K) contiguous; rebuilt whenever the reference changes
        self._h0_flat = self.h0_sparse.reshape(3, -1).contiguous()

        f_dense_h = self.f_dense.cpu().numpy()
        inds = (f_dense_h >= freqs[0]) & (f_dense_h <= freqs[-1])
        inds_d = torch.as_tensor(inds, device=self.device)
        self.f_dense = self.f_dense[inds_d].contiguous()
        self.d = self.d[:, inds_d].contiguous()
        h0 = h0[:, inds_d]
        f_n_host = f_dense_h[inds]

        freqs_d = to_dev(freqs, self.device)
        f_m = (freqs_d[1:] + freqs_d[:-1]) / 2
        df = float(f_n_host[1] - f_n_host[0])

        self.sens_mat.update_frequency_arr(f_n_host)
        S_n = to_dev(np.asarray([self.sens_mat[0], self.sens_mat[1], self.sens_mat[2]]), self.device)

        # A0, A1, B0, B1 per segment and channel, and the band totals: one
        # deterministic segmented reduction on the device (likelihood.py:499-533)
        K = self.length_f_het
        self.data_constants = torch.zeros(12 * K, dtype=torch.complex128, device=self.device)
        totals = torch.zeros(3, dtype=torch.float64, device=self.device)
        self.backend.hdyn_prep(self.data_constants, totals, self.f_dense, self.d, h0.contiguous(),
                               S_n.contiguous(), freqs_d, df)
        self.reference_d_d, self.reference_h_h, self.reference_d_h = (float(v) for v in totals.cpu().numpy())
        self.reference_ll = -1 / 2 * (self.reference_d_d + self.reference_h_h - 2 * self.reference_d_h)

        self.freqs = freqs_d
        self.f_m = f_m
        template_gen_kwargs["squeeze"] = False
        self.template_gen_kwargs = template_gen_kwargs

    def _merge_kwargs(self, waveform_kwargs):
        all_keys = list(set(list(waveform_kwargs.keys()) + list(self.template_gen_kwargs.keys())))
        for key in all_keys:
            if key in ["direct", "compress", "squeeze"]:
                waveform_kwargs[key] = self.template_gen_kwargs[key]
            elif key in self.template_gen_kwargs:
                waveform_kwargs[key] = waveform_kwargs.get(key, self.template_gen_kwargs[key])
        waveform_kwargs["freqs"] = self.freqs
        return waveform_kwargs

    def get_ll(self, params, return_extracted_snr=False, phase_marginalize=False, fused=True, **waveform_kwargs):
        """log-likelihood of each parameter column (likelihood.py:550-648).
        ``params``: (12,) or (12, num_bin_all); returns numpy (num_bin_all,)."""
        torch = self.xp
        self.phase_marginalize = phase_marginalize
        self.return_extracted_snr = return_extracted_snr
        waveform_kwargs = self._merge_kwargs(dict(waveform_kwargs))
        if fused and not (phase_marginalize or return_extracted_snr):
            # the logL the evaluation kernel writes (the same expression of its
            # d_h, h_h), one device-to-host copy
            rows, t_start, t_end = self.fused_inputs(params, waveform_kwargs)
            return self.get_ll_device(rows, t_start, t_end, modes=waveform_kwargs.get("modes")).cpu().numpy()
        if fused:
            self.hdyn_d_h, self.hdyn_h_h = self._fused_inner_products(params, waveform_kwargs)
        else:
            self.h_sparse = self.template_gen(*params, **waveform_kwargs)
            r = self.h_sparse / self.h0_sparse
            B = self.template_gen.num_bin_all
            self.hdyn_d_h = torch.zeros(B, dtype=torch.complex128, device=self.device)
            self.hdyn_h_h = torch.zeros(B, dtype=torch.complex128, device=self.device)
            residuals_in = r.permute(2, 1, 0).contiguous().flatten()
            self.like_gen(self.hdyn_d_h, self.hdyn_h_h, residuals_in, self.data_constants, self.freqs, B,
                          int(self.freqs.numel()), 3, self.include_last_segment)
        d_h_temp = self.hdyn_d_h if not self.phase_marginalize else torch.abs(self.hdyn_d_h)
        out = (-1 / 2.0 * (self.reference_d_d + self.hdyn_h_h - 2 * d_h_temp).real).cpu().numpy()
        if self.return_extracted_snr:
            snr = (d_h_temp.real / torch.sqrt(self.hdyn_h_h.real)).cpu().numpy()
            return np.array([out, snr]).T
        return out

    # ---------------------------------------------------------------- fused path

    def fused_inputs(self, params, waveform_kwargs=None):
        """Device inputs of the fused kernel: (11, B) parameters (m1 >= m2
        enforced as phenomhm.py:175-199 does), t_start, t_end."""
        waveform_kwargs = {} if waveform_kwargs is None else waveform_kwargs
        return fused_param_rows(self.template_gen, params, waveform_kwargs.get("t_obs_start", 0.0),
                                waveform_kwargs.get("t_obs_end", 1.0), self.device)

    def get_ll_device(self, rows, t_start, t_end, precision=0, modes=None):
        """Fused log-likelihood of device-resident inputs (from ``fused_inputs``);
        returns a device tensor. No host transfer."""
        ll = self.xp.empty(rows.shape[1], dtype=self.xp.float64, device=self.device)
        d_h, h_h = self._fused_core(rows, t_start, t_end, modes, precision, ll)
        self.hdyn_d_h, self.hdyn_h_h = d_h, h_h
        return ll

    def _fused_core(self, rows, t_start, t_end, modes, precision, ll=None):
        torch = self.xp
        gen = self.template_gen
        if modes is None:
            modes = self.template_gen_kwargs.get("modes") or gen.amp_phase_gen.allowable_modes
        ells = np.asarray([l for l, m in modes], dtype=np.int32)
        mms = np.asarray([m for l, m in modes], dtype=np.int32)
        B = rows.shape[1]
        d_h = torch.empty(B, dtype=torch.complex128, device=self.device)
        h_h = torch.empty(B, dtype=torch.complex128, device=self.device)
        resp = gen.response_gen
        self.backend.hdyn_fused(d_h, h_h, rows, t_start, t_end, ells, mms, self.freqs, self._h0_flat,
                                self.data_constants, gen.amp_phase_gen.qnm_packed, gen.amp_phase_gen.dspin,
                                gen.amp_phase_gen.initial_t_val, resp.cpp_response, resp.TDItag_int, resp.rescaled,
                                resp.tdi2, self.include_last_segment, precision, logL=ll,
                                d_d=self.reference_d_d)
        gen.num_bin_all = B
        return d_h, h_h

    def _fused_inner_products(self, params, waveform_kwargs, precision=0):
        rows, t_start, t_end = self.fused_inputs(params, waveform_kwargs)
        return self._fused_core(rows, t_start, t_end, waveform_kwargs.get("modes"), precision)

    # ---------------------------------------------------------------- gradient

    GRAD_COORDS = ("lnM", "q", "chi1z", "chi2z", "distance", "phi_ref", "cos_inc", "lam", "sin_beta", "psi", "t_ref")

    @staticmethod
    def sampling_to_params(x):
        """(B, 11) sampling coordinates (ln(m1 + m2), q = m1/m2 >= 1, chi1z,
        chi2z, distance [m], phi_ref, cos inc, lam, sin beta, psi, t_ref [s]) ->
        the (12, B) parameter rows of ``get_ll`` (f_ref = 0)."""
        x = np.atleast_2d(np.asarray(to_host(x), dtype=np.float64))
        if x.shape[1] != 11:
            raise ValueError("x must be (B, 11)")
        M = np.exp(x[:, 0])
        q = x[:, 1]
        out = np.zeros((12, x.shape[0]))
        out[0] = M * q / (1 + q)
        out[1] = M / (1 + q)
        out[2], out[3], out[4], out[5] = x[:, 2], x[:, 3], x[:, 4], x[:, 5]
        out[7] = np.arccos(x[:, 6])
        out[8] = x[:, 7]
        out[9] = np.arcsin(x[:, 8])
        out[10], out[11] = x[:, 9], x[:, 10]
        return out

    def get_ll_grad_device(self, rows, t_start, t_end, modes=None):
        """log-likelihood (B,) and its gradient (11, B) in the sampling
        coordinates, from device inputs of ``fused_inputs``; device tensors."""
        torch = self.xp
        gen = self.template_gen
        if modes is None:
            modes = self.template_gen_kwargs.get("modes") or gen.amp_phase_gen.allowable_modes
        ells = np.asarray([l for l, m in modes], dtype=np.int32)
        mms = np.asarray([m for l, m in modes], dtype=np.int32)
        B = rows.shape[1]
        d_h = torch.empty(B, dtype=torch.complex128, device=self.device)
        h_h = torch.empty(B, dtype=torch.complex128, device=self.device)
        grad = torch.empty((11, B), dtype=torch.float64, device=self.device)
        ll = torch.empty(B, dtype=torch.float64, device=self.device)
        resp = gen.response_gen
        self.backend.hdyn_fused_grad(d_h, h_h, grad, rows, t_start, t_end, ells, mms, self.freqs, self._h0_flat,
                                     self.data_constants, gen.amp_phase_gen.qnm_packed, gen.amp_phase_gen.dspin,
                                     gen.amp_phase_gen.initial_t_val, resp.cpp_response, resp.TDItag_int,
                                     resp.rescaled, resp.tdi2, self.include_last_segment, logL=ll,
                                     d_d=self.reference_d_d)
        gen.num_bin_all = B
        self.hdyn_d_h, self.hdyn_h_h = d_h, h_h
        return ll, grad

    def get_ll_and_grad(self, x, **waveform_kwargs):
        """log-likelihood (B,) and gradient (B, 11) at sampling coordinates x
        (B, 11) (see ``sampling_to_params``); numpy arrays."""
        waveform_kwargs = self._merge_kwargs(dict(waveform_kwargs))
        rows, t_start, t_end = self.fused_inputs(self.sampling_to_params(x), waveform_kwargs)
        ll, grad = self.get_ll_grad_device(rows, t_start, t_end, waveform_kwargs.get("modes"))
        return ll.cpu().numpy(), grad.T.cpu().numpy()

    def loglike(self, x):
        """Differentiable log-likelihood: x is a (B, 11) float64 torch tensor of
        sampling coordinates; returns (B,) with a backward pass through the
        adjoint kernel (``HeterodynedLogLike``)."""
        return HeterodynedLogLike.apply(x, self)


class HeterodynedLogLike(torch.autograd.Function):
    """torch.autograd.Function over the fused heterodyned log-likelihood:
    forward runs the gradient kernel (value and gradient in one pass) and keeps
    the (B, 11) gradient for backward. Arguments: x (B, 11) sampling
    coordinates (float64 tensor), the HeterodynedLikelihood. For x on the GPU
    everything stays on the device (``sampling_rows_device``): no host round
    trip, capturable in a graph."""

    @staticmethod
    def forward(ctx, x, like):
        xd = x.detach().to(device=like.device, dtype=torch.float64)
        ll, grad = like.get_ll_grad_device(*sampling_rows_device(xd))
        ctx.save_for_backward(grad.T.to(x.device, x.dtype))
        return ll.to(x.device, x.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out[:, None] * g, None


class DirectLogLike(torch.autograd.Function):
    """torch.autograd.Function over the fused direct log-likelihood of a ragged
    batch: forward runs the adjoint launch sequence (value and gradient in one
    pass) and keeps the (B, 11) gradient for backward. Arguments: x (B, 11)
    sampling coordinates (float64 tensor), the Likelihood, lengths, plan, modes.
    For x on the GPU and a plan given, everything stays on the device; the
    window-search flags are not read back (``get_ll_and_grad_fused`` checks them)."""

    @staticmethod
    def forward(ctx, x, like, lengths, plan, modes):
        xd = x.detach().to(device=like.device, dtype=torch.float64)
        if plan is None:
            plan = like.direct_plan(np.asarray(lengths, dtype=np.int32), modes)
        flags = torch.zeros(xd.shape[0], dtype=torch.int32, device=like.device)
        ll, grad = like.get_ll_grad_fused_device(plan, *sampling_rows_device(xd), flags, modes)
        ctx.save_for_backward(grad.T.to(x.device, x.dtype))
        return ll.to(x.device, x.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out[:, None] * g, None, None, None, None
