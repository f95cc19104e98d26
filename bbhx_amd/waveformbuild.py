"""Template generation: PhenomHM -> response -> (spline ->) TDI channels.

Host-side mirror of the reference's ``waveformbuild.py`` (``TemplateInterpFD``
:36-239 and ``BBHWaveformFD`` :242-700): same constructor and call
signatures, keyword semantics and return shapes. Device arrays are torch
tensors; all per-binary kernels run as one launch each.
"""

from __future__ import annotations

import warnings

import numpy as np

from .response.fastfdresponse import LISATDIResponse
from .utils.arrays import as_np1d, to_dev
from .utils.constants import YRSID_SI
from .utils.interpolate import CubicSplineInterpolant
from .utils.parallelbase import BBHxParallelModule
from .utils.transform import tSSBfromLframe
from .waveforms.phenomhm import PhenomHMAmpPhase


class TemplateInterpFD(BBHxParallelModule):
    """Evaluate the sparse-grid splines on the dense data grid for every binary
    (waveformbuild.py:36-239)."""

    def __init__(self, force_backend=None):
        super().__init__(force_backend=force_backend)

    @property
    def template_gen(self):
        return self.backend.InterpTDI_wrap

    @property
    def template_channels(self):
        return [self.template_carrier[i].reshape(self.num_channels, int(self.lengths[i]))
                for i in range(self.num_bin_all)]

    def band_indices(self, data_freqs, freqs_shaped):
        """Per binary: the dense band [start, end) covered by its sparse grids
        (searchsorted side='right' of the min/max sparse frequency,
        waveformbuild.py:139-156) and, per mode and dense bin, the spline segment
        searchsorted(x_bm, f, 'right') - 1 (:159-172)."""
        torch = self.xp
        B, M = freqs_shaped.shape[:2]
        lo = freqs_shaped.amin(dim=(1, 2))
        hi = freqs_shaped.amax(dim=(1, 2))
        st = torch.searchsorted(data_freqs, lo, right=True)
        et = torch.searchsorted(data_freqs, hi, right=True)
        st_h = st.cpu().numpy().astype(np.int64)
        et_h = et.cpu().numpy().astype(np.int64)
        lengths = (et_h - st_h).astype(np.int32)
        inds = []
        for i in range(B):
            seg = data_freqs[st_h[i]:et_h[i]]
            vals = seg[None, :].expand(M, -1).contiguous()
            k = torch.searchsorted(freqs_shaped[i].contiguous(), vals, right=True) - 1
            inds.append(k.to(torch.int32).flatten().contiguous())
        return st_h.astype(np.int32), lengths, inds

    def __call__(self, data_freqs, interp_container, t_start, t_end, length, num_modes, num_channels):
        """Returns the list of (num_channels, n_b) complex templates (waveformbuild.py:89-239)."""
        torch = self.xp
        num_bin_all = len(np.atleast_1d(t_start))
        self.length = length
        self.num_modes = num_modes
        self.num_bin_all = num_bin_all
        data_freqs = to_dev(data_freqs, self.device)
        self.data_length = int(data_freqs.numel())
        self.num_channels = num_channels
        self.t_start, self.t_end = as_np1d(t_start), as_np1d(t_end)
        (freqs, y, c1, c2, c3) = interp_container
        freqs_shaped = freqs.reshape(num_bin_all, num_modes, -1)

        self.start_inds, self.lengths, inds = self.band_indices(data_freqs, freqs_shaped)
        self.inds = inds
        self.ptrs = np.asarray([a.data_ptr() for a in inds], dtype=np.int64)
        total = int(np.sum(self.lengths.astype(np.int64))) * num_channels
        self._template_flat = torch.zeros(max(total, 1), dtype=torch.complex128, device=self.device)
        self.template_carrier = []
        off = 0
        for n in self.lengths:
            self.template_carrier.append(self._template_flat[off:off + num_channels * int(n)])
            off += num_channels * int(n)
        template_carrier_ptrs = np.asarray([t.data_ptr() for t in self.template_carrier], dtype=np.int64)
        self.template_gen(template_carrier_ptrs, data_freqs, freqs, y, c1, c2, c3, as_np1d(t_start),
                          as_np1d(t_end), self.length, self.data_length, self.num_bin_all, self.num_modes, self.ptrs,
                          self.start_inds, self.lengths)
        return self.template_channels


class BBHWaveformFD(BBHxParallelModule):
    """Binary-black-hole TDI waveforms (waveformbuild.py:242-700).

    Args:
        amp_phase_kwargs (dict): for :class:`PhenomHMAmpPhase`.
        response_kwargs (dict): for :class:`LISATDIResponse`.
        interp_kwargs (dict): for :class:`TemplateInterpFD`.
        force_backend: backend name or instance.
    """

    def __init__(self, amp_phase_kwargs=None, response_kwargs=None, interp_kwargs=None, force_backend=None):
        amp_phase_kwargs = {} if amp_phase_kwargs is None else amp_phase_kwargs
        response_kwargs = {} if response_kwargs is None else response_kwargs
        interp_kwargs = {} if interp_kwargs is None else interp_kwargs
        super().__init__(force_backend=force_backend)
        self.force_backend = self.backend
        self.amp_phase_gen = PhenomHMAmpPhase(**amp_phase_kwargs, force_backend=self.backend)
        self.response_gen = LISATDIResponse(**response_kwargs, force_backend=self.backend)
        self.num_interp_params = 9
        self.interp_response = TemplateInterpFD(**interp_kwargs, force_backend=self.backend)

    @property
    def waveform_gen(self):
        return self.backend.direct_sum_wrap

    def time_window(self, lam, beta, t_obs_start, t_obs_end):
        """SSB-frame window of each binary (waveformbuild.py:477-484)."""
        if isinstance(t_obs_start, np.ndarray) is False and not np.isscalar(t_obs_start):
            raise ValueError("t_obs_start and t_obs_end must be floats or numpy arrays even when running on GPUs.")
        if isinstance(t_obs_end, np.ndarray) is False and not np.isscalar(t_obs_end):
            raise ValueError("t_obs_start and t_obs_end must be floats or numpy arrays even when running on GPUs.")
        t_start = np.atleast_1d(tSSBfromLframe(t_obs_start * YRSID_SI, lam, beta, 0.0))
        t_end = np.atleast_1d(tSSBfromLframe(t_obs_end * YRSID_SI, lam, beta, 0.0))
        return t_start, t_end

    def __call__(self, m1, m2, chi1z, chi2z, distance, phi_ref, f_ref, inc, lam, beta, psi, t_ref,
                 t_obs_start=0.0, t_obs_end=1.0, freqs=None, length=None, modes=None, shift_t_limits=True,
                 direct=False, compress=True, squeeze=False, fill=False, combine=False):
        torch = self.xp
        m1, m2, chi1z, chi2z = as_np1d(m1), as_np1d(m2), as_np1d(chi1z), as_np1d(chi2z)
        distance, phi_ref, inc = as_np1d(distance), as_np1d(phi_ref), as_np1d(inc)
        lam, beta, psi, t_ref = as_np1d(lam), as_np1d(beta), as_np1d(psi), as_np1d(t_ref)
        self.num_bin_all = len(m1)
        if shift_t_limits is False:
            warnings.warn("Deprecated: shift_t_limits. This option is removed and permanently set to "
                          "shift_t_limits=True.")
        t_start, t_end = self.time_window(lam, beta, t_obs_start, t_obs_end)

        if freqs is None and length is None:
            raise ValueError("Must input freqs or length.")
        elif freqs is not None and direct is True:
            length = len(freqs) if freqs.ndim == 1 else freqs.shape[1]
        elif direct is False:
            self.data_length = len(freqs)
            if length is None:
                raise ValueError("If direct is False, length parameter must be given.")
        self.length = length

        if modes is None:
            self.num_modes = len(self.amp_phase_gen.allowable_modes)
        else:
            if not isinstance(modes, list):
                raise ValueError("modes must be a list.")
            self.num_modes = len(modes)

        out_buffer = torch.zeros(self.num_interp_params * self.length * self.num_modes * self.num_bin_all,
                                 dtype=torch.float64, device=self.device)
        freqs_temp = freqs if direct else None
        phi_ref_amp_phase = np.zeros_like(m1)
        Tobs = t_end - t_start
        self.amp_phase_gen(m1, m2, chi1z, chi2z, distance, phi_ref_amp_phase, f_ref, t_ref, length, freqs=freqs_temp,
                           out_buffer=out_buffer, modes=modes, Tobs=Tobs, direct=direct)

        self.response_gen(self.amp_phase_gen.freqs, inc, lam, beta, psi, phi_ref, length, out_buffer=out_buffer,
                          modes=self.amp_phase_gen.modes, direct=direct)

        self.out_buffer_final = out_buffer.reshape(9, self.num_bin_all, self.num_modes, self.length)
        t_start_d, t_end_d = to_dev(t_start, self.device), to_dev(t_end, self.device)

        if direct and compress:
            templateChannels = torch.zeros(self.num_bin_all * 3 * self.length, dtype=torch.complex128,
                                           device=self.device)
            self.waveform_gen(templateChannels, out_buffer, self.num_bin_all, self.length, 3, self.num_modes,
                              t_start_d, t_end_d)
            out = templateChannels.reshape(self.num_bin_all, 3, self.length)
            return out.squeeze() if squeeze else out

        elif direct:
            out = torch.zeros((self.num_bin_all, 3, self.num_modes, self.length), dtype=torch.complex128,
                              device=self.device)
            for mode_i in range(self.num_modes):
                templateChannels = torch.zeros(self.num_bin_all * 3 * self.length, dtype=torch.complex128,
                                               device=self.device)
                out_buffer_temp = self.out_buffer_final[:, :, mode_i, :].contiguous().flatten()
                self.waveform_gen(templateChannels, out_buffer_temp, self.num_bin_all, self.length, 3, 1, t_start_d,
                                  t_end_d)
                out[:, :, mode_i, :] += templateChannels.reshape(self.num_bin_all, 3, self.length)
            return out.squeeze() if squeeze else out

        spline = CubicSplineInterpolant(self.amp_phase_gen.freqs, out_buffer, length=self.length,
                                        num_interp_params=self.num_interp_params, num_modes=self.num_modes,
                                        num_bin_all=self.num_bin_all, force_backend=self.backend)
        self.spline = spline
        template_channels = self.interp_response(freqs, spline.container, t_start, t_end, self.length,
                                                 self.num_modes, 3)
        if fill:
            nf = len(freqs)
            if combine:
                data_out = torch.zeros((3, nf), dtype=torch.complex128, device=self.device)
                for temp, start_i, length_i in zip(template_channels, self.interp_response.start_inds,
                                                   self.interp_response.lengths):
                    data_out[:, int(start_i):int(start_i) + int(length_i)] += temp
                return data_out.squeeze() if squeeze else data_out
            data_out = torch.zeros((self.num_bin_all, 3, nf), dtype=torch.complex128, device=self.device)
            for bin_i, (temp, start_i, length_i) in enumerate(zip(template_channels, self.interp_response.start_inds,
                                                                  self.interp_response.lengths)):
                data_out[bin_i, :, int(start_i):int(start_i) + int(length_i)] = temp
            return data_out
        return (template_channels, self.interp_response.start_inds, self.interp_response.lengths)
