"""PhenomHM / PhenomD in the amplitude-phase representation on the GPU.

Host-side mirror of the reference's ``waveforms/phenomhm.py`` (class
``PhenomHMAmpPhase``, :32-614) with the same constructor, call signature,
flat buffer layout and properties. Device arrays are torch tensors; the
kernels behind ``self.backend`` are the gfx950 ones (``include/bbhx_hip.h``).
"""

from __future__ import annotations

import numpy as np
from scipy.interpolate import CubicSpline

from ..utils.arrays import as_np1d, np_logspace_rows, to_dev
from ..utils.constants import MSUN_SI, MTSUN_SI
from ..utils.parallelbase import BBHxParallelModule
from .ringdownphenomd import QNMData_a, QNMData_fdamp, QNMData_fring, dspin


class PhenomHMAmpPhase(BBHxParallelModule):
    """PhenomHM (or PhenomD) amplitude, phase and t(f) (phenomhm.py:32-614).

    Args:
        force_backend: backend name or instance (default ``"hip"``).
        run_phenomd (bool): run PhenomD, i.e. the (2,2) mode only.
        mf_min, mf_max (float): dimensionless frequency range of the default grid.
        initial_t_val (float): start of the time window; shifts the phase only.
    """

    def __init__(self, force_backend=None, run_phenomd=False, mf_min=1e-4, mf_max=0.6, initial_t_val=0.0):
        super().__init__(force_backend=force_backend)
        torch = self.xp
        self.run_phenomd = run_phenomd
        self.allowable_modes = [(2, 2), (3, 3), (4, 4), (2, 1), (3, 2), (4, 3)]
        self.ells_default = torch.tensor([2, 3, 4, 2, 3, 4], dtype=torch.int32, device=self.device)
        self.mms_default = torch.tensor([2, 3, 4, 1, 2, 3], dtype=torch.int32, device=self.device)
        if self.run_phenomd:
            self.allowable_modes = [(2, 2)]
            self.ells_default = torch.tensor([2], dtype=torch.int32, device=self.device)
            self.mms_default = torch.tensor([2], dtype=torch.int32, device=self.device)
        self.mf_min = mf_min
        self.mf_max = mf_max
        self.initial_t_val = initial_t_val
        self._init_phenomd_fring_spline()

    @property
    def waveform_gen(self):
        return self.backend.waveform_amp_phase_wrap

    @property
    def phenomhm_ringdown_freqs(self):
        return self.backend.get_phenomhm_ringdown_frequencies

    @property
    def phenomd_ringdown_freqs(self):
        return self.backend.get_phenomd_ringdown_frequencies

    def _init_phenomd_fring_spline(self):
        """Cubic-spline coefficients of the QNM table (phenomhm.py:148-163). The
        spline has 1000 segments; the coefficient arrays are padded to the 1001
        table entries so every index the kernel can form stays inside them."""
        spl_ring = CubicSpline(QNMData_a, QNMData_fring)
        spl_damp = CubicSpline(QNMData_a, QNMData_fdamp)

        def pad(c):
            out = np.zeros(QNMData_a.size)
            out[: c.size] = c
            return out

        host = [QNMData_fring, pad(spl_ring.c[-2]), pad(spl_ring.c[-3]), pad(spl_ring.c[-4]),
                QNMData_fdamp, pad(spl_damp.c[-2]), pad(spl_damp.c[-3]), pad(spl_damp.c[-4])]
        self.qnm_packed = to_dev(np.concatenate(host), self.device)
        n = QNMData_a.size
        (self.y_rd, self.c1_rd, self.c2_rd, self.c3_rd,
         self.y_dm, self.c1_dm, self.c2_dm, self.c3_dm) = [self.qnm_packed[i * n:(i + 1) * n] for i in range(8)]
        self.dspin = dspin

    def _sanity_check_modes(self, ells, mms):
        for ell, mm in zip(ells, mms):
            if (int(ell), int(mm)) not in self.allowable_modes:
                raise ValueError(
                    "Requested mode [(l,m) = ({},{})] is not available. Allowable modes include {}".format(
                        ell, mm, self.allowable_modes))

    def _sanity_check_params(self, m1, m2, chi1z, chi2z):
        """|chi| <= 1 and m1 >= m2 with the spins swapped along (phenomhm.py:175-199)."""
        if np.any(np.abs(chi1z) > 1.0):
            raise ValueError("chi1z array contains values with abs(chi1z) > 1.0 which is not allowed.")
        if np.any(np.abs(chi2z) > 1.0):
            raise ValueError("chi2z array contains values with abs(chi2z) > 1.0 which is not allowed.")
        for name, a in (("m1", m1), ("m2", m2), ("chi1z", chi1z), ("chi2z", chi2z)):
            if not np.all(np.isfinite(a)):
                raise ValueError(f"{name} contains non-finite values.")
        if np.any(m1 <= 0.0) or np.any(m2 <= 0.0):
            raise ValueError("masses must be positive.")
        switch = m1 < m2
        m1_temp = m2[switch]
        m2[switch] = m1[switch]
        m1[switch] = m1_temp
        chi1z_temp = chi2z[switch]
        chi2z[switch] = chi1z[switch]
        chi1z[switch] = chi1z_temp
        return m1, m2, chi1z, chi2z

    def _initialize_waveform_container(self):
        self.waveform_carrier = self.xp.zeros(self.nparams * self.length * self.num_modes * self.num_bin_all,
                                              dtype=self.xp.float64, device=self.device)

    def _initialize_freqs(self, m1, m2, mms):
        """Default grid Mf in [mf_min, mf_max] * m/2 (phenomhm.py:208-224)."""
        torch = self.xp
        M_tot_sec = to_dev((m1 + m2) * MTSUN_SI, self.device)
        lo = torch.tensor(np.log10(self.mf_min), dtype=torch.float64, device=self.device)
        hi = torch.tensor(np.log10(self.mf_max), dtype=torch.float64, device=self.device)
        base = np_logspace_rows(lo, hi, self.length)
        self.freqs = (base[None, None, :] / M_tot_sec[:, None, None] * mms.to(torch.float64)[None, :, None]
                      / 2).flatten()

    @property
    def amp(self):
        return self.waveform_carrier[0: self.num_per_param].reshape(self.num_bin_all, self.num_modes, self.length)

    @property
    def phase(self):
        return self.waveform_carrier[self.num_per_param: 2 * self.num_per_param].reshape(
            self.num_bin_all, self.num_modes, self.length)

    @property
    def tf(self):
        return self.waveform_carrier[2 * self.num_per_param: 3 * self.num_per_param].reshape(
            self.num_bin_all, self.num_modes, self.length)

    @property
    def freqs_shaped(self):
        return self._freqs.reshape(self.num_bin_all, self.num_modes, self.length)

    @property
    def freqs(self):
        return self._freqs

    @freqs.setter
    def freqs(self, f):
        self._freqs = f.flatten() if f.ndim > 1 else f

    def _modes_to_arrays(self, modes):
        torch = self.xp
        if modes is not None:
            ells = np.asarray([ell for ell, mm in modes], dtype=np.int32)
            mms = np.asarray([mm for ell, mm in modes], dtype=np.int32)
            self.modes = modes
            self._sanity_check_modes(ells, mms)
            return (torch.as_tensor(ells, device=self.device), torch.as_tensor(mms, device=self.device))
        self.modes = self.allowable_modes
        return self.ells_default, self.mms_default

    def run_wave(self, m1, m2, chi1z, chi2z, distance, phi_ref, f_ref, t_ref, length, freqs=None,
                 out_buffer=None, modes=None):
        """Generate amplitude, phase and t(f) (phenomhm.py:270-507)."""
        torch = self.xp
        m1, m2, chi1z, chi2z = as_np1d(m1), as_np1d(m2), as_np1d(chi1z), as_np1d(chi2z)
        distance, phi_ref, f_ref, t_ref = as_np1d(distance), as_np1d(phi_ref), as_np1d(f_ref), as_np1d(t_ref)
        m1, m2, chi1z, chi2z = self._sanity_check_params(m1, m2, chi1z, chi2z)

        ells, mms = self._modes_to_arrays(modes)
        if self.run_phenomd:
            ells = torch.tensor([2], dtype=torch.int32, device=self.device)
            mms = torch.tensor([2], dtype=torch.int32, device=self.device)
            self.modes = self.allowable_modes

        num_modes = int(ells.numel())
        num_bin_all = len(m1)
        self.length = int(length)
        self.num_modes = num_modes
        self.num_bin_all = num_bin_all
        self.nparams = 3
        self.num_per_param = self.length * num_modes * num_bin_all
        self.num_per_bin = self.length * num_modes

        dev = self.device
        m1_d, m2_d = to_dev(m1, dev), to_dev(m2, dev)
        chi1_d, chi2_d = to_dev(chi1z, dev), to_dev(chi2z, dev)
        dist_d, fref_d = to_dev(distance, dev), to_dev(f_ref, dev)

        if out_buffer is None:
            self._initialize_waveform_container()
        else:
            if out_buffer.numel() < 3 * self.num_per_param:
                raise ValueError("out_buffer is too short for (3, num_bin_all, num_modes, length).")
            self.waveform_carrier = out_buffer

        if freqs is None:
            self._initialize_freqs(m1, m2, mms)
        else:
            freqs = to_dev(freqs, dev)
            if freqs.ndim == 1:
                self.freqs = freqs.reshape(1, 1, -1).expand(num_bin_all, num_modes, -1).contiguous().flatten()
            elif freqs.ndim == 3:
                self.freqs = freqs.flatten().clone()
            else:
                raise ValueError("If providing freqs, must be 1D array (length,) or 3D array "
                                 "(num_bin_all, num_modes, length).")
        if self._freqs.numel() != self.num_per_param:
            raise ValueError("freqs do not match (num_bin_all, num_modes, length).")

        m1_SI, m2_SI = m1_d * MSUN_SI, m2_d * MSUN_SI

        fringdown = torch.zeros(num_modes * num_bin_all, dtype=torch.float64, device=dev)
        fdamp = torch.zeros_like(fringdown)
        self.phenomd_ringdown_freqs(fringdown, fdamp, m1_d, m2_d, chi1_d, chi2_d, num_bin_all, self.y_rd,
                                    self.c1_rd, self.c2_rd, self.c3_rd, self.y_dm, self.c1_dm, self.c2_dm,
                                    self.c3_dm, self.dspin)
        if not self.run_phenomd:
            append_frd = fringdown[:num_bin_all].clone()
            append_fdm = fdamp[:num_bin_all].clone()
            self.phenomhm_ringdown_freqs(fringdown, fdamp, m1_d, m2_d, chi1_d, chi2_d, ells, mms, num_modes,
                                         num_bin_all)
            fringdown = torch.cat([fringdown.reshape(-1, num_modes), append_frd[:, None]], dim=1).flatten().contiguous()
            fdamp = torch.cat([fdamp.reshape(-1, num_modes), append_fdm[:, None]], dim=1).flatten().contiguous()
        self.fringdown, self.fdamp = fringdown, fdamp

        # inside the kernel t_ref and phi_ref are zero
        self.waveform_gen(self.waveform_carrier, ells.contiguous(), mms.contiguous(), self._freqs, m1_SI, m2_SI,
                          chi1_d, chi2_d, dist_d, fref_d, num_modes, self.length, num_bin_all, fringdown, fdamp,
                          self.run_phenomd)

        # phase += 2 pi f (t_ref - t0); tf += t_ref (phenomhm.py:487-507)
        tref_d = to_dev(t_ref, dev)
        temp = self.freqs_shaped * (tref_d[:, None, None] - self.initial_t_val) * 2 * np.pi
        npp = self.num_per_param
        self.waveform_carrier[npp: 2 * npp] += temp.flatten()
        self.waveform_carrier[2 * npp: 3 * npp] += tref_d[:, None, None].expand(
            num_bin_all, num_modes, self.length).flatten()

    @staticmethod
    def tobs_grid(Tobs, B, M, device):
        """The observation length per (binary, mode) as a (B, M) float64 tensor:
        None (no upper cut), a float, a (B,) or a (B, M) numpy array
        (phenomhm.py:561-575)."""
        import torch

        if Tobs is None:
            return torch.full((B, M), 1e300, dtype=torch.float64, device=device)
        if isinstance(Tobs, float):
            return torch.full((B, M), Tobs, dtype=torch.float64, device=device)
        if isinstance(Tobs, np.ndarray) and Tobs.ndim == 1 and Tobs.shape[0] == B:
            return to_dev(Tobs, device)[:, None].expand(B, M).contiguous()
        if isinstance(Tobs, np.ndarray) and Tobs.shape == (B, M):
            return to_dev(Tobs, device)
        raise ValueError("Tobs entered incorrectly.")

    @staticmethod
    def window_bounds(tf, fs, Tobs_t):
        """Frequency range of each (binary, mode) grid with 0 <= t(f) <= Tobs
        (phenomhm.py:520-600): where t(f) crosses 0 (or Tobs) inside the grid,
        the crossing frequency by linear interpolation between the two grid
        points around it; otherwise the grid's own end. tf, fs: (B, M, L);
        Tobs_t: (B, M). Raises if a grid crosses a bound more than once."""
        import torch

        def cross(level, lower):
            f_end = (fs.min(dim=-1).values if lower else fs.max(dim=-1).values).clone()
            fix = (tf.min(dim=-1).values < level) if lower else (tf.max(dim=-1).values > level)
            nfix = int(fix.sum().item())
            if nfix == 0:
                return f_end
            tfx, fsx, lv = tf[fix], fs[fix], level[fix][:, None]
            hit = torch.nonzero((lv < tfx[:, 1:]) & (lv > tfx[:, :-1]))
            if hit.shape[0] != nfix:
                where = "x_min" if lower else "x_max"
                hint = "zero" if lower else "the observation time"
                raise ValueError(f"The search for a window at {where} did not run correctly. If t_ref is close to "
                                 f"{hint}, the time-frequency correspondence may be crossing zero twice.")
            seg = hit[:, 1]
            rows = torch.arange(nfix, device=tf.device)
            y1, y2 = fsx[rows, seg], fsx[rows, seg + 1]
            x1, x2 = tfx[rows, seg], tfx[rows, seg + 1]
            f_end[fix] = (y2 - y1) / (x2 - x1) * (lv[:, 0] - x1) + y1
            return f_end

        zero = torch.zeros_like(Tobs_t)
        return cross(zero, True), cross(Tobs_t, False)

    def __call__(self, *args, Tobs=None, direct=False, **kwargs):
        """Run the waveform; unless ``direct``, refit each (binary, mode) grid to
        the part with 0 <= t(f) <= Tobs and run again (phenomhm.py:509-614)."""
        torch = self.xp
        self.run_wave(*args, **kwargs)
        if direct:
            return
        Tobs_t = self.tobs_grid(Tobs, self.num_bin_all, self.num_modes, self.device)
        f_min, f_max = self.window_bounds(self.tf, self.freqs_shaped, Tobs_t)
        new_freqs = np_logspace_rows(torch.log10(f_min), torch.log10(f_max), self.length)
        tmp_kwargs = kwargs.copy()
        tmp_kwargs["freqs"] = new_freqs
        self.run_wave(*args, **tmp_kwargs)
