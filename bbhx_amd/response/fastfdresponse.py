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

    def __call__(self, freqs, inc, lam, beta, psi, phi_ref, length, modes=None, phase=None, tf=None,
                 out_buffer=None, adjust_phase=True, direct=False):
        """Evaluate the response (fastfdresponse.py:278-531)."""
        torch = self.xp
        dev = self.device
        inc, lam, beta, psi, phi_ref = (to_dev(as_np1d(v), dev) for v in (inc, lam, beta, psi, phi_ref))

        if modes is not None:
            ells_h = np.asarray([ell for ell, mm in modes], dtype=np.int32)
            mms_h = np.asarray([mm for ell, mm in modes], dtype=np.int32)
            self._sanity_check_modes(ells_h, mms_h)
            ells = torch.as_tensor(ells_h, device=dev)
            mms = torch.as_tensor(mms_h, device=dev)
        else:
            ells, mms = self.ells_default, self.mms_default
        self.modes = [(int(ell), int(mm)) for ell, mm in zip(ells.cpu().tolist(), mms.cpu().tolist())]

        num_modes = int(ells.numel())
        num_bin_all = int(inc.numel())
        self.length = int(length)
        self.num_modes = num_modes
        self.num_bin_all = num_bin_all
        self.num_per_param = self.length * num_modes * num_bin_all
        self.num_per_bin = self.length * num_modes
        self.nresponse_params = 6

        if out_buffer is None and phase is None and tf is None:
            raise ValueError("Must provide either out_buffer or both phase and tf.")

        if out_buffer is None:
            includes_amps = 0
            self.nparams = includes_amps + 2 + self.nresponse_params
            self._initialize_response_container()
        else:
            nparams_empirical = out_buffer.numel() / (self.num_bin_all * self.num_modes * self.length)
            if np.allclose(nparams_empirical, np.round(nparams_empirical)) and int(nparams_empirical) in [8, 9]:
                self.nparams = int(nparams_empirical)
            else:
                raise ValueError(
                    f"out_buffer incorrect length. The length should be equivalent to (8 or 9) * "
                    f"{self.num_bin_all * self.num_modes * self.length}. Given length is {out_buffer.numel()}.")
            includes_amps = 1 if self.nparams == 9 else 0
            self.response_carrier = out_buffer
        self.includes_amps = includes_amps

        freqs = to_dev(freqs, dev)
        if freqs.ndim > 1:
            if tuple(freqs.shape) != (self.num_bin_all, self.num_modes, self.length):
                raise ValueError(f"freqs have incorrect shape. Shape should be "
                                 f"{(self.num_bin_all, self.num_modes, self.length)}. Current shape is {tuple(freqs.shape)}.")
            freqs = freqs.flatten()
        elif freqs.numel() != self.num_bin_all * self.num_modes * self.length:
            raise ValueError(f"freqs incorrect length. The length should be equivalent to "
                             f"{self.num_bin_all * self.num_modes * self.length}. Given length is {freqs.numel()}.")
        self.freqs = freqs

        t_lo, t_hi = float(self.orbits.t_base.min()), float(self.orbits.t_base.max())
        n = self.num_per_param
        if phase is not None and tf is not None:
            use_phase_tf = True
            if tuple(phase.shape) != tuple(tf.shape):
                raise ValueError("Shape of phase array and tf array need to be the same shape.")
            phase_d, tf_d = to_dev(phase, dev), to_dev(tf, dev)
            if not direct and (float(tf_d.min()) < t_lo or float(tf_d.max()) > t_hi):
                raise ValueError(f"Orbital information does not cover minimum ({float(tf_d.min())}) and maximum "
                                 f"({float(tf_d.max())}) tf. Orbital information begins at {t_lo} and ends at {t_hi}.")
            if phase_d.numel() != n:
                raise ValueError(f"phase incorrect length. The length should be equivalent to {n}. "
                                 f"Given length is {phase_d.numel()}.")
            self.response_carrier[includes_amps * n:(includes_amps + 1) * n] = phase_d.flatten()
            self.response_carrier[(includes_amps + 1) * n:(includes_amps + 2) * n] = tf_d.flatten()
        elif phase is not None or tf is not None:
            raise ValueError("If provided phase or tf, need to provide both.")
        else:
            use_phase_tf = False
            if not direct:
                tfv = self.tf
                lo, hi = float(tfv.min()), float(tfv.max())
                if lo < t_lo or hi > t_hi:
                    raise ValueError(f"Orbital information does not cover minimum ({lo}) and maximum ({hi}) tf. "
                                     f"Orbital information begins at {t_lo} and ends at {t_hi}.")

        self.response_gen(self.response_carrier, ells, mms, self.freqs, phi_ref, inc, lam, beta, psi,
                          self.TDItag_int, self.rescaled, self.tdi2, self.order_fresnel_stencil, num_modes,
                          self.length, num_bin_all, includes_amps)

        if use_phase_tf and adjust_phase:
            output = self.response_carrier[includes_amps * n:(includes_amps + 1) * n]
            if isinstance(phase, torch.Tensor):
                phase.copy_(output.reshape(phase.shape).to(phase.device))
            else:
                phase[...] = output.reshape(phase.shape).cpu().numpy()
