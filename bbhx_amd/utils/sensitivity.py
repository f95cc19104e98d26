"""Analytic first-generation TDI A/E/T noise PSDs.

The reference takes its noise model from lisaanalysistools
(``AET1SensitivityMatrix``, ``get_sensitivity(..., sens_fn="A1TDISens")``;
likelihood.py:23, :357-363, :493-496; tests/test_bbhx.py:224-226), which is not
part of the reference tree. This module implements the standard LISA
requirement noise model (test-mass acceleration 3e-15 m s^-2 Hz^-1/2 with the
0.4 mHz / 8 mHz shaping, optical-metrology 15 pm Hz^-1/2 with the 2 mHz
shaping; Babak, Hewitson & Petiteau, arXiv:2108.01167) and the equal-arm
TDI-1 A/E/T transfer of those noises:

  S_A = S_E = 8 sin^2 x [ 2 S_pm (3 + 2 cos x + cos 2x) + S_op (2 + cos x) ]
  S_T = 16 S_op (1 - cos x) sin^2 x + 128 S_pm sin^2 x sin^4(x/2),   x = 2 pi f L / c

in fractional-frequency units. The PSD is an input shared by the GPU path and
the CPU oracle; no parity claim depends on its exact values.
"""

from __future__ import annotations

import numpy as np

from .constants import C_SI

LISA_L = 2.5e9


def lisa_noises(f, acc=3e-15, oms=15e-12):
    f = np.asarray(f, dtype=np.float64)
    two_pi_f = 2.0 * np.pi * f
    s_acc = acc**2 * (1.0 + (0.4e-3 / f) ** 2) * (1.0 + (f / 8e-3) ** 4)  # (m/s^2)^2/Hz
    spm = s_acc / two_pi_f**4 * (two_pi_f / C_SI) ** 2                     # relative frequency
    sop = oms**2 * (1.0 + (2e-3 / f) ** 4) * (two_pi_f / C_SI) ** 2
    return spm, sop


def A1TDISens(f, **kwargs):
    x = 2.0 * np.pi * LISA_L / C_SI * np.asarray(f, dtype=np.float64)
    spm, sop = lisa_noises(f, **kwargs)
    return 8.0 * np.sin(x) ** 2 * (2.0 * spm * (3.0 + 2.0 * np.cos(x) + np.cos(2 * x)) + sop * (2.0 + np.cos(x)))


E1TDISens = A1TDISens


def T1TDISens(f, **kwargs):
    x = 2.0 * np.pi * LISA_L / C_SI * np.asarray(f, dtype=np.float64)
    spm, sop = lisa_noises(f, **kwargs)
    return 16.0 * sop * (1.0 - np.cos(x)) * np.sin(x) ** 2 + 128.0 * spm * np.sin(x) ** 2 * np.sin(0.5 * x) ** 4


_FUNCS = {"A1TDISens": A1TDISens, "E1TDISens": E1TDISens, "T1TDISens": T1TDISens}


def get_sensitivity(f, sens_fn="A1TDISens", **kwargs):
    """PSD of one channel at frequencies f (same call shape as lisaanalysistools')."""
    return _FUNCS[sens_fn](f, **kwargs)


class AET1SensitivityMatrix:
    """A, E, T PSDs on a frequency grid; ``update_frequency_arr`` and indexing as
    the reference uses them (likelihood.py:493-496)."""

    def __init__(self, f):
        self.update_frequency_arr(f)

    def update_frequency_arr(self, f):
        self.frequency_arr = np.asarray(f, dtype=np.float64)
        self._s = [A1TDISens(self.frequency_arr), E1TDISens(self.frequency_arr), T1TDISens(self.frequency_arr)]

    def __getitem__(self, i):
        return self._s[i]

    def __len__(self):
        return 3
