"""Time and sky-frame conversions (the reference's ``utils/transform.py``,
after Marsat et al., arXiv:2003.00357): SSB <-> LISA-frame time
(:75-136), SSB <-> LISA-frame sky position and polarisation (:136-310) and
the eight LISA-frame sky modes of an MBHB (:313-426)."""

from __future__ import annotations

import numpy as np

from .constants import AU_SI, C_SI, YRSID_SI, ConstOmega

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


# ------------------------------------------------------------------ sky frames
#
# The LISA frame (arXiv:2003.00357, Sec. II) is the SSB ecliptic frame turned
# by zeta = 60 degrees about the horizontal axis perpendicular to the Sun
# direction alpha = Omega t: unit vectors map as
#     k_L = Rz(alpha) Ry(zeta) Rz(-alpha) k_SSB,
# so the LISA-frame pole seen from the SSB frame is
#     n(alpha) = (-sin(zeta) cos(alpha), -sin(zeta) sin(alpha), cos(zeta)),
# and sin(beta_L) = k_SSB . n(alpha). The polarisation angle changes by the
# angle between the two frames' meridians at the source.

ZETA = np.pi / 3.0


def modpi(phase):
    """phase modulo pi, in [0, pi) (transform.py:43-56)."""
    return phase - np.floor(phase / np.pi) * np.pi


def mod2pi(phase):
    """phase modulo 2 pi, in [0, 2 pi) (transform.py:59-72)."""
    return phase - np.floor(phase / (2 * np.pi)) * 2 * np.pi


def _unit(lam, beta):
    cb = np.cos(beta)
    return np.stack([cb * np.cos(lam), cb * np.sin(lam), np.sin(beta) * np.ones_like(lam)])


def _turn(k, alpha, zeta):
    """Rz(alpha) Ry(zeta) Rz(-alpha) applied to unit vectors k (3, ...)."""
    ca, sa = np.cos(alpha), np.sin(alpha)
    cz, sz = np.cos(zeta), np.sin(zeta)
    x = ca * k[0] + sa * k[1]  # Rz(-alpha)
    y = -sa * k[0] + ca * k[1]
    z = k[2]
    x, z = cz * x + sz * z, -sz * x + cz * z  # Ry(zeta)
    return np.stack([ca * x - sa * y, sa * x + ca * y, z])  # Rz(alpha)


def SSB_to_LISA(tSSB, lambdaSSB, betaSSB, psiSSB, t0=0.0):
    """Time, ecliptic longitude and latitude and polarisation in the LISA
    frame from those in the SSB frame (transform.py:227-310); ``t0`` in years.
    The phase is not transformed (approximant dependent)."""
    alpha = ConstOmega * tSSB + ConstOmega * (t0 * YRSID_SI)
    tL = tLfromSSBframe(tSSB, lambdaSSB, betaSSB, t0 * YRSID_SI)
    kL = _turn(_unit(lambdaSSB, betaSSB), alpha, ZETA)
    lambdaL = np.arctan2(kL[1], kL[0])
    betaL = np.arcsin(kL[2])
    d = lambdaSSB - alpha
    psiL = modpi(psiSSB + np.arctan2(-np.sin(ZETA) * np.sin(d),
                                     np.cos(betaSSB) * np.cos(ZETA) + np.sin(ZETA) * np.sin(betaSSB) * np.cos(d)))
    return tL, lambdaL, betaL, psiL


def LISA_to_SSB(tL, lambdaL, betaL, psiL, t0=0.0):
    """Time, ecliptic longitude and latitude and polarisation in the SSB frame
    from those in the LISA frame (transform.py:136-224); ``t0`` in years. The
    Sun direction alpha depends on the SSB time, which itself depends on the
    SSB sky position: both are refined together three times, starting from
    t_SSB = t_L, as the reference does."""
    t0s = t0 * YRSID_SI
    kL = _unit(lambdaL, betaL)
    tSSB = tL
    for _ in range(3):
        alpha = ConstOmega * tSSB + ConstOmega * t0s
        k = _turn(kL, alpha, -ZETA)
        lambdaSSB = np.arctan2(k[1], k[0])
        betaSSB = np.arcsin(k[2])
        tSSB = tSSBfromLframe(tL, lambdaSSB, betaSSB, t0s)
    d = lambdaL - alpha
    psiSSB = modpi(psiL + np.arctan2(np.sin(ZETA) * np.sin(d),
                                     np.cos(betaL) * np.cos(ZETA) - np.sin(ZETA) * np.sin(betaL) * np.cos(d)))
    return tSSB, lambdaSSB % (2 * np.pi), betaSSB, psiSSB


def mbh_sky_mode_transform(coords, ind_map=None, kind="both", inplace=False, cos_i=False):
    """The 8 LISA-frame sky modes of an MBHB (transform.py:313-426): the two
    latitudinal modes (inc, beta, psi) -> (pi - inc or -cos inc, -beta, pi - psi)
    and the four longitudinal ones lam + k pi/2 (mod 2 pi), psi + k pi/2 (mod pi).

    coords (num, ndim); ind_map gives the columns of inc, lam, beta, psi
    (default 7, 8, 9, 10). Not in place: coords is tiled factor times (8 for
    "both", 4 for "long", 2 for "lat") and row r gets the mode the reference
    assigns to it (latitudinal flip on odd rows; longitudinal shift k on rows
    r = k mod 4 for "long", r = 2k or 2k + 1 mod 8 for "both"). In place: the
    row count must be a multiple of the factor."""
    if ind_map is None:
        ind_map = dict(inc=7, lam=8, beta=9, psi=10)
    elif not isinstance(ind_map, dict):
        raise ValueError("If providing the ind_map kwarg, it must be a dict.")
    factors = {"both": 8, "long": 4, "lat": 2}
    if kind not in factors:
        raise ValueError("The kwarg 'kind' must be lat for latitudinal transformation, long for longitudinal "
                         "transformation, or both for both.")
    factor = factors[kind]
    if inplace:
        if coords.shape[0] % factor != 0:
            raise ValueError(f"If performing an inplace transformation, the coords provided must have a first "
                             f"dimension size divisible by {factor} for a '{kind}' transformation.")
    else:
        coords = np.tile(coords, (factor, 1))
    inc, lam, beta, psi = ind_map["inc"], ind_map["lam"], ind_map["beta"], ind_map["psi"]
    rows = np.arange(coords.shape[0])
    if kind in ("both", "lat"):
        odd = rows % 2 == 1
        coords[odd, inc] = -coords[odd, inc] if cos_i else np.pi - coords[odd, inc]
        coords[odd, beta] = -coords[odd, beta]
        coords[odd, psi] = np.pi - coords[odd, psi]
    if kind in ("both", "long"):
        shift = (rows % 4) if kind == "long" else ((rows % 8) // 2)
        sel = shift > 0
        coords[sel, lam] = (coords[sel, lam] + np.pi / 2 * shift[sel]) % (2 * np.pi)
        coords[sel, psi] = (coords[sel, psi] + np.pi / 2 * shift[sel]) % np.pi
    return coords
