"""Equal-arm LISA orbits, tabulated in the lisaanalysistools ``Orbits`` layout.

The reference takes orbits from lisaanalysistools (``response/fastfdresponse.py:23,
:102-175``): spacecraft positions ``x`` (N, 3, 3), link unit vectors ``n``
(N, 6, 3) and light travel times ``L`` (N, 6) on a uniform time grid, linearly
interpolated in time by the response (``configure(linear_interp_setup=True)``).
That package and its orbit files are not part of the reference tree, so this
module provides its own equal-arm model in the same layout:

* a rigid, cartwheeling triangle on heliocentric orbits, to first order in the
  eccentricity e = L / (2 sqrt(3) R) (Rubbo, Cornish & Poujade 2004, eqs. 56-58;
  R = 1 AU, L = 2.5e9 m, one orbit per sidereal year, phases alpha = Omega t,
  beta_n = 2 pi (n - 1) / 3);
* positions are absolute (solar-system-barycentre frame), as in
  lisaanalysistools orbit files;
* link ``ij`` is received by spacecraft i and emitted by j; its unit vector
  points from j to i, ``n_ij = (x_i - x_j) / |x_i - x_j|``.

Any object exposing the same attributes (``pycppdetector_args``, ``t_base``,
``LINKS``) can be passed to :class:`LISATDIResponse` instead.
"""

from __future__ import annotations

import numpy as np

from .constants import AU_SI, C_SI, YRSID_SI

LINKS = [12, 23, 31, 13, 32, 21]
SC_R = [1, 2, 3, 1, 3, 2]
SC_E = [2, 3, 1, 3, 2, 1]


def equal_arm_positions(t, armlength=2.5e9, radius=AU_SI, period=YRSID_SI):
    """Spacecraft positions (len(t), 3, 3) [m] of the first-order equal-arm model."""
    t = np.atleast_1d(np.asarray(t, dtype=np.float64))
    e = armlength / (2.0 * np.sqrt(3.0) * radius)
    alpha = 2.0 * np.pi / period * t
    sa, ca = np.sin(alpha), np.cos(alpha)
    out = np.zeros((t.size, 3, 3))
    for n in range(3):
        b = 2.0 * np.pi * n / 3.0
        sb, cb = np.sin(b), np.cos(b)
        out[:, n, 0] = radius * ca + radius * e * (sa * ca * sb - (1.0 + sa * sa) * cb)
        out[:, n, 1] = radius * sa + radius * e * (sa * ca * cb - (1.0 + ca * ca) * sb)
        out[:, n, 2] = -np.sqrt(3.0) * radius * e * np.cos(alpha - b)
    return out


class EqualArmlengthOrbits:
    """Tabulated equal-arm orbits (stand-in for lisaanalysistools'
    ``EqualArmlengthOrbits``; same attribute names and argument tuple).

    Args:
        dt (float): table step [s] (default 1e4).
        t_max (float): table span [s] from t = 0 (default 10 sidereal years).
        armlength (float): nominal arm length [m] used by the response (default 2.5e9).
        force_backend: backend name or instance for the device copies.
    """

    LINKS = LINKS
    SC_R = SC_R
    SC_E = SC_E

    def __init__(self, dt=1e4, t_max=10.0 * YRSID_SI, armlength=2.5e9, force_backend=None):
        self.dt = float(dt)
        self.N = int(np.ceil(t_max / self.dt)) + 1
        self.armlength = float(armlength)
        self.t_base = np.arange(self.N) * self.dt
        self.x_base = equal_arm_positions(self.t_base, armlength=armlength)
        n = np.zeros((self.N, 6, 3))
        ltt = np.zeros((self.N, 6))
        for k, (r, s) in enumerate(zip(SC_R, SC_E)):
            d = self.x_base[:, r - 1] - self.x_base[:, s - 1]
            norm = np.sqrt((d * d).sum(axis=-1))
            n[:, k] = d / norm[:, None]
            ltt[:, k] = norm / C_SI
        self.n_base = n
        self.ltt_base = ltt
        self._force_backend = force_backend
        self._dev = None

    def configure(self, linear_interp_setup=True, **kwargs):
        """The tables already sit on a uniform grid; kept for interface parity."""
        return self

    def to_device(self, device):
        """Device copies of (n, L, x) as contiguous float64 tensors."""
        import torch

        if self._dev is None or self._dev[0].device != torch.device(device):
            self._dev = tuple(torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)
                              for a in (self.n_base, self.ltt_base, self.x_base))
        return self._dev

    def pycppdetector_args_on(self, device):
        """Arguments of ``add_orbit_information`` (lisaresponse.pyx:41-71) with device tables."""
        n, L, x = self.to_device(device)
        return (self.dt, self.N, n, L, x, np.asarray(LINKS, np.int32), np.asarray(SC_R, np.int32),
                np.asarray(SC_E, np.int32), self.armlength)

    def as_host_dict(self):
        """Tables for the CPU oracle."""
        return dict(dt=self.dt, N=self.N, n_arr=self.n_base, x_arr=self.x_base, links=np.asarray(LINKS, np.int32),
                    armlength=self.armlength)
