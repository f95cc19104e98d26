"""LISA fast frequency-domain TDI response on the GPU.

Host-side mirror of the reference's ``response/fastfdresponse.py``
(``LISATDIResponse``, :26-531): same constructor, call signature, buffer
checks and properties; the kernel is the gfx950 one behind the stateful
``pyFastLISAResponse`` object of the backend.
"""

from __future__ import annotations

import numpy as np

from ..utils.arrays import as_np1d, to_dev
from ..utils.orbits import EqualArmlengthOrbits
from ..utils.parallelbase import BBHxParallelModule


class LISATDIResponse(BBHxParallelModule):
    """Fast frequency-domain TDI response (arXiv:1806.10734, arXiv:2003.00357).

    Args:
        TDItag (str): ``"XYZ"`` or ``"AET"`` (default).
        orbits: orbit tables (default :class:`EqualArmlengthOrbits`).
        rescaled (bool): rescale the TDI transfer functions.
        tdi2 (bool): multiply TDI-1 by -2i sin(4x) e^{4ix} (AET only, as in the reference).
        order_fresnel_stencil (int): must be 0.
        force_backend: backend name or instance.
    """

    def __init__(self, TDItag="AET", orbits=None, rescaled=False, tdi2=False, order_fresnel_stencil=0,
                 force_backend=None):
        self.rescaled = rescaled
        self.tdi2 = tdi2
        super().__init__(force_backend=force_backend)
        if order_fresnel_stencil > 0:
            raise NotImplementedError
        self.order_fresnel_stencil = order_fresnel_stencil
        self.TDItag = TDItag
        self.TDItag_int = 1 if TDItag == "XYZ" else 2
        torch = self.xp
        self.allowable_modes = [(2, 2), (3, 3), (4, 4), (2, 1), (3, 2), (4, 3)]
        self.ells_default = torch.tensor([2, 3, 4, 2, 3, 4], dtype=torch.int32, device=self.device)
        self.mms_default = torch.tensor([2, 3, 4, 1, 2, 3], dtype=torch.int32, device=self.device)
        self.orbits = orbits
        self.cpp_response = self.backend.pyFastLISAResponse()
        self.cpp_response.add_orbit_information(*self.check_add_orbit_args(*self._orbit_args()))

    def _orbit_args(self):
        if hasattr(self.orbits, "pycppdetector_args_on"):
            return self.orbits.pycppdetector_args_on(self.device)
        args = list(self.orbits.pycppdetector_args)
        for k in (2, 3, 4):
            args[k] = to_dev(args[k], self.device)
        return tuple(args)

    def check_add_orbit_args(self, *args):
        """Arguments must be [dt, N, n_arr, L_arr, x_arr, links, sc_r, sc_e, armlength]
        (fastfdresponse.py:107-134)."""
        import torch

        try:
            assert len(args) == 9
            assert isinstance(args[0], float)
            assert isinstance(args[1], int)
            for k in (2, 3, 4):
                assert isinstance(args[k], torch.Tensor) and args[k].dtype == torch.float64
            assert len(args[5]) == len(self.orbits.LINKS)
            assert len(args[6]) == len(self.orbits.LINKS)
            assert len(args[7]) == len(self.orbits.LINKS)
            assert isinstance(args[8], float)
        except AssertionError:
            raise ValueError("Arguments for cpp class are not correct.")
        return args

    @property
    def cpp_response(self):
        if self._cpp_response is None:
            raise ValueError("Must add cpp_response and add orbit information.")
        return self._cpp_response

    @cpp_response.setter
    def cpp_response(self, cpp_response):
        self._cpp_response = cpp_response

    @property
    def response_gen(self):
        return self.cpp_response.LISA_response_wrap

    @property
    def orbits(self):
        return self._orbits

    @orbits.setter
    def orbits(self, orbits):
        if orbits is None:
            self._orbits = EqualArmlengthOrbits()
        elif not (hasattr(orbits, "pycppdetector_args") or hasattr(orbits, "pycppdetector_args_on")):
            raise ValueError("Input orbits must provide the Orbits table interface (pycppdetector_args).")
        else:
            self._orbits = orbits
        self._orbits.configure(linear_interp_setup=True)

    def _sanity_check_modes(self, ells, mms):
        for ell, mm in zip(ells, mms):
            if (int(ell), int(mm)) not in self.allowable_modes:
                raise ValueError(
                    "Requested mode [(l,m) = ({},{})] is not available. Allowable modes include {}".format(
                        ell, mm, self.allowable_modes))

    def _initialize_response_container(self):
        self.response_carrier = self.xp.zeros(self.nparams * self.length * self.num_modes * self.num_bin_all,
                                              dtype=self.xp.float64, device=self.device)

    def _slab(self, k):
        n = self.num_per_param
        return self.response_carrier[(self.includes_amps + k) * n:(self.includes_amps + k + 1) * n].reshape(
            self.num_bin_all, self.num_modes, self.length)

    @property
    def transferL1(self):
        return self._slab(2) + 1j * self._slab(3)

    @property
    def transferL2(self):
        return self._slab(4) + 1j * self._slab(5)

    @property
    def transferL3(self):
        return self._slab(6) + 1j * self._slab(7)

    @property
    def phase(self):
        return self._slab(0)

    @property
    def tf(self):
        return self._slab(1)

    # ---- host-side checks (no device work; the reference's raise points)

    @staticmethod
    def buffer_layout(numel, num_bin_all, num_modes, length):
        """(nparams, includes_amps) of a caller's out_buffer of ``numel`` doubles:
        8 slabs (phase, tf, 6 transfer parts) or 9 (amp first) of B * M * L each
        (fastfdresponse.py:389-421); anything else is a ValueError."""
        per = num_bin_all * num_modes * length
        slabs, rem = divmod(int(numel), per) if per > 0 else (0, 1)
        if rem or slabs not in (8, 9):
            raise ValueError(f"out_buffer incorrect length. The length should be equivalent to (8 or 9) * {per}. "
                             f"Given length is {int(numel)}.")
        return slabs, int(slabs == 9)

    @staticmethod
    def check_freqs(shape, numel, num_bin_all, num_modes, length):
        """Frequencies are either (B, M, L) or flat with B * M * L entries."""
        want = (num_bin_all, num_modes, length)
        if len(shape) > 1:
            if tuple(shape) != want:
                raise ValueError(f"freqs have incorrect shape. Shape should be {want}. Current shape is {tuple(shape)}.")
        elif int(numel) != num_bin_all * num_modes * length:
            raise ValueError(f"freqs incorrect length. The length should be equivalent to "
                             f"{num_bin_all * num_modes * length}. Given length is {int(numel)}.")

    @staticmethod
    def check_orbit_coverage(tf_min, tf_max, t_lo, t_hi):
        """Every t(f) must lie inside the orbit table (fastfdresponse.py:486-492,
        where the reference stops in a debugger; here it is a ValueError)."""
        if tf_min < t_lo or tf_max > t_hi:
            raise ValueError(f"Orbital information does not cover minimum ({tf_min}) and maximum ({tf_max}) tf. "
                             f"Orbital information begins at {t_lo} and ends at {t_hi}.")

    def _mode_arrays(self, modes):
        torch = self.xp
        if modes is None:
            return self.ells_default, self.mms_default
        ells_h = np.asarray([ell for ell, mm in modes], dtype=np.int32)
        mms_h = np.asarray([mm for ell, mm in modes], dtype=np.int32)
        self._sanity_check_modes(ells_h, mms_h)
        return torch.as_tensor(ells_h, device=self.device), torch.as_tensor(mms_h, device=self.device)

    def __call__(self, freqs, inc, lam, beta, psi, phi_ref, length, modes=None, phase=None, tf=None,
                 out_buffer=None, adjust_phase=True, direct=False):
        """Transfer functions and Doppler-shifted phase of every (binary, mode,
        frequency) (fastfdresponse.py:278-531). Either ``out_buffer`` (the
        waveform buffer, 8 or 9 slabs, phase and tf already in place) or both
        ``phase`` and ``tf`` are given; with the latter the shifted phase is
        written back into ``phase`` when ``adjust_phase``."""
        dev = self.device
        inc, lam, beta, psi, phi_ref = (to_dev(as_np1d(v), dev) for v in (inc, lam, beta, psi, phi_ref))
        ells, mms = self._mode_arrays(modes)
        self.modes = [(int(ell), int(mm)) for ell, mm in zip(ells.cpu().tolist(), mms.cpu().tolist())]
        B, M, L = int(inc.numel()), int(ells.numel()), int(length)
        self.length, self.num_modes, self.num_bin_all = L, M, B
        self.num_per_bin = L * M
        self.num_per_param = n = B * M * L
        self.nresponse_params = 6

        given_pt = phase is not None or tf is not None
        if given_pt and (phase is None or tf is None):
            raise ValueError("If provided phase or tf, need to provide both.")
        if out_buffer is None and not given_pt:
            raise ValueError("Must provide either out_buffer or both phase and tf.")
        if out_buffer is None:
            self.nparams, self.includes_amps = 8, 0
            self._initialize_response_container()
        else:
            self.nparams, self.includes_amps = self.buffer_layout(out_buffer.numel(), B, M, L)
            self.response_carrier = out_buffer
        ia = self.includes_amps

        freqs = to_dev(freqs, dev)
        self.check_freqs(tuple(freqs.shape), freqs.numel(), B, M, L)
        self.freqs = freqs.flatten() if freqs.ndim > 1 else freqs

        t_lo, t_hi = float(self.orbits.t_base.min()), float(self.orbits.t_base.max())
        if given_pt:
            if tuple(phase.shape) != tuple(tf.shape):
                raise ValueError("Shape of phase array and tf array need to be the same shape.")
            phase_d, tf_d = to_dev(phase, dev), to_dev(tf, dev)
            if phase_d.numel() != n:
                raise ValueError(f"phase incorrect length. The length should be equivalent to {n}. "
                                 f"Given length is {phase_d.numel()}.")
            if not direct:
                self.check_orbit_coverage(float(tf_d.min()), float(tf_d.max()), t_lo, t_hi)
            self.response_carrier[ia * n:(ia + 1) * n] = phase_d.flatten()
            self.response_carrier[(ia + 1) * n:(ia + 2) * n] = tf_d.flatten()
        elif not direct:
            tfv = self.tf
            self.check_orbit_coverage(float(tfv.min()), float(tfv.max()), t_lo, t_hi)

        self.response_gen(self.response_carrier, ells, mms, self.freqs, phi_ref, inc, lam, beta, psi,
                          self.TDItag_int, self.rescaled, self.tdi2, self.order_fresnel_stencil, M, L, B, ia)

        if given_pt and adjust_phase:
            shifted = self.response_carrier[ia * n:(ia + 1) * n].reshape(phase.shape)
            if isinstance(phase, self.xp.Tensor):
                phase.copy_(shifted.to(phase.device))
            else:
                phase[...] = shifted.cpu().numpy()
