"""Time-frame conversions on the likelihood path (the reference's
``utils/transform.py:75-136``, after Marsat et al., arXiv:2003.00357)."""

from __future__ import annotations

import numpy as np

from .constants import AU_SI, C_SI, ConstOmega

OrbitR = AU_SI


def tSSBfromLframe(tL, lambdaSSB, betaSSB, t0=0.0):
    """Solar-system-barycentre time from the retarded time at the LISA
    constellation centre (transform.py:75-105)."""
    ConstPhi0 = ConstOmega * t0
    phase = ConstOmega * tL + ConstPhi0 - lambdaSSB
    RoC = OrbitR / C_SI
    return (tL + RoC * np.cos(betaSSB) * np.cos(phase)
            - 1.0 / 2 * ConstOmega * pow(RoC * np.cos(betaSSB), 2) * np.sin(2.0 * phase))


def tLfromSSBframe(tSSB, lambdaSSB, betaSSB, t0=0.0):
    """Retarded time at the constellation centre from the SSB time, to first
    order in R/c as the reference does (transform.py:108-133)."""
    ConstPhi0 = ConstOmega * t0
    phase = ConstOmega * tSSB + ConstPhi0 - lambdaSSB
    RoC = OrbitR / C_SI
    return tSSB - RoC * np.cos(betaSSB) * np.cos(phase)
