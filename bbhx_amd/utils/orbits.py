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

Orbits from elsewhere (a lisaanalysistools orbit file exported to arrays, a
numerical orbit) enter through :class:`TabulatedOrbits`, which takes the
arrays of ``add_orbit_information`` (lisaresponse.pyx:41-71) directly or from
an ``.npz`` file. Any object exposing the same attributes
(``pycppdetector_args_on`` or ``pycppdetector_args``, ``t_base``, ``LINKS``)
can be passed to :class:`LISATDIResponse` as well.
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
        return (self.dt, self.N, n, L, x, np.asarray(self.LINKS, np.int32), np.asarray(self.SC_R, np.int32),
                np.asarray(self.SC_E, np.int32), self.armlength)

    def as_host_dict(self):
        """Tables for the CPU oracle."""
        return dict(dt=self.dt, N=self.N, n_arr=self.n_base, x_arr=self.x_base, links=np.asarray(self.LINKS, np.int32),
                    armlength=self.armlength)

    def to_arrays(self):
        """The tables as the keyword arguments of :class:`TabulatedOrbits`."""
        return dict(dt=self.dt, n_arr=self.n_base, L_arr=self.ltt_base, x_arr=self.x_base,
                    links=np.asarray(self.LINKS, np.int32), sc_r=np.asarray(self.SC_R, np.int32),
                    sc_e=np.asarray(self.SC_E, np.int32), armlength=self.armlength)


class TabulatedOrbits(EqualArmlengthOrbits):
    """Orbit tables given as arrays, in the layout of the reference's
    ``add_orbit_information(dt, N, n_arr, L_arr, x_arr, links, sc_r, sc_e,
    armlength)`` (lisaresponse.pyx:41-71): link unit vectors n (N, 6, 3), light
    travel times L (N, 6) [s], spacecraft positions x (N, 3, 3) [m] on the
    uniform grid t = k dt, k = 0..N-1; ``links`` the six link ids (ij: received
    by spacecraft i, emitted by j) with their receiver ``sc_r`` and emitter
    ``sc_e`` spacecraft. The arrays are checked (shapes, the six ids 12, 23,
    31, 13, 32, 21 once each, sc_r / sc_e consistent with them, unit vectors,
    dt > 0) and then used as given.

    ``TabulatedOrbits.from_file(path)`` reads an ``.npz`` archive with the keys
    dt, n, L, x, links, sc_r, sc_e and armlength (lisaanalysistools' HDF5
    orbit files need h5py, which is not part of this build: export their
    datasets to such an archive)."""

    def __init__(self, dt, n_arr, L_arr, x_arr, links=LINKS, sc_r=SC_R, sc_e=SC_E, armlength=2.5e9,
                 force_backend=None):
        n = np.ascontiguousarray(np.asarray(n_arr, dtype=np.float64))
        L = np.ascontiguousarray(np.asarray(L_arr, dtype=np.float64))
        x = np.ascontiguousarray(np.asarray(x_arr, dtype=np.float64))
        links = [int(v) for v in np.asarray(links).ravel()]
        sc_r = [int(v) for v in np.asarray(sc_r).ravel()]
        sc_e = [int(v) for v in np.asarray(sc_e).ravel()]
        if not (float(dt) > 0.0):
            raise ValueError("dt must be positive")
        N = x.shape[0] if x.ndim == 3 else -1
        if N < 2 or x.shape != (N, 3, 3) or n.shape != (N, 6, 3) or L.shape != (N, 6):
            raise ValueError(f"orbit arrays must be x (N, 3, 3), n (N, 6, 3), L (N, 6) with N >= 2; got "
                             f"{x.shape}, {n.shape}, {L.shape}")
        if sorted(links) != sorted(LINKS) or len(sc_r) != 6 or len(sc_e) != 6:
            raise ValueError(f"links must hold each of {LINKS} once, with six sc_r and sc_e entries")
        for l, r, e in zip(links, sc_r, sc_e):
            if (r, e) != (l // 10, l % 10):
                raise ValueError(f"link {l} must be received by spacecraft {l // 10} and emitted by {l % 10}")
        if not np.all(np.isfinite(x)) or not np.all(np.isfinite(n)) or not np.all(np.isfinite(L)):
            raise ValueError("orbit arrays must be finite")
        if np.max(np.abs(np.sqrt((n * n).sum(-1)) - 1.0)) > 1e-6:
            raise ValueError("link vectors n must be unit vectors")
        if not (float(armlength) > 0.0):
            raise ValueError("armlength must be positive")
        self.LINKS, self.SC_R, self.SC_E = links, sc_r, sc_e
        self.dt = float(dt)
        self.N = N
        self.armlength = float(armlength)
        self.t_base = np.arange(N) * self.dt
        self.x_base, self.n_base, self.ltt_base = x, n, L
        self._force_backend = force_backend
        self._dev = None

    @classmethod
    def from_file(cls, path, **kwargs):
        with np.load(path, allow_pickle=False) as z:
            return cls(float(z["dt"]), z["n"], z["L"], z["x"], links=z["links"], sc_r=z["sc_r"], sc_e=z["sc_e"],
                       armlength=float(z["armlength"]), **kwargs)

    def save(self, path):
        """Write the tables to an ``.npz`` archive that ``from_file`` reads."""
        np.savez(path, dt=self.dt, n=self.n_base, L=self.ltt_base, x=self.x_base, links=np.asarray(self.LINKS),
                 sc_r=np.asarray(self.SC_R), sc_e=np.asarray(self.SC_E), armlength=self.armlength)
