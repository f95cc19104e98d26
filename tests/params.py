"""Seeded parameter draws of the benchmark configurations (shared by tests and bench).

Distributions: total mass log-uniform, q uniform [1, 10], chi1, chi2 uniform
[-0.99, 0.99], distance uniform 10-100 Gpc, phi_ref U[0, 2pi], cos(inc)
U[-1, 1], lam U[0, 2pi], sin(beta) U[-1, 1], psi U[0, pi], t_ref
U[0.1, 0.9] yr, f_ref = 0. Rows: m1, m2, chi1z, chi2z, distance, phi_ref,
f_ref, inc, lam, beta, psi, t_ref.
"""

from __future__ import annotations

import numpy as np

MODES = [(2, 2), (2, 1), (3, 3), (3, 2), (4, 4), (4, 3)]
YRSID_SI = 31558149.763545603
PC_SI = 3.0856775814913674e+16

PRIOR_RANGES = {
    "lnM": (np.log(1e5), np.log(1e7)), "q": (1.0, 10.0), "chi1": (-0.99, 0.99), "chi2": (-0.99, 0.99),
    "dist": (10e9 * PC_SI, 100e9 * PC_SI), "phi": (0.0, 2 * np.pi), "cosi": (-1.0, 1.0), "lam": (0.0, 2 * np.pi),
    "sinb": (-1.0, 1.0), "psi": (0.0, np.pi), "tref": (0.1 * YRSID_SI, 0.9 * YRSID_SI),
}


def draw(n, seed, mtot=(1e5, 1e7), tref=(0.1, 0.9)):
    rng = np.random.default_rng(seed)
    M = np.exp(rng.uniform(np.log(mtot[0]), np.log(mtot[1]), n))
    q = rng.uniform(1.0, 10.0, n)
    m1 = M * q / (1 + q)
    m2 = M / (1 + q)
    chi1 = rng.uniform(-0.99, 0.99, n)
    chi2 = rng.uniform(-0.99, 0.99, n)
    dist = rng.uniform(10, 100, n) * 1e9 * PC_SI
    phi = rng.uniform(0, 2 * np.pi, n)
    inc = np.arccos(rng.uniform(-1, 1, n))
    lam = rng.uniform(0, 2 * np.pi, n)
    beta = np.arcsin(rng.uniform(-1, 1, n))
    psi = rng.uniform(0, np.pi, n)
    t_ref = rng.uniform(tref[0], tref[1], n) * YRSID_SI
    return np.stack([m1, m2, chi1, chi2, dist, phi, np.zeros(n), inc, lam, beta, psi, t_ref])


def perturb(ref, n, seed, scale=1e-3):
    """Templates around ``ref``: each of the 11 sampled parameters offset by
    N(0, scale * prior range) in its sampling coordinate."""
    rng = np.random.default_rng(seed)
    m1, m2, c1, c2, d, phi, fref, inc, lam, beta, psi, tref = ref
    lnM = np.log(m1 + m2) + scale * np.diff(PRIOR_RANGES["lnM"])[0] * rng.standard_normal(n)
    q = m1 / m2 + scale * 9.0 * rng.standard_normal(n)
    q = np.maximum(q, 1.0)
    M = np.exp(lnM)
    out = np.zeros((12, n))
    out[0] = M * q / (1 + q)
    out[1] = M / (1 + q)
    out[2] = np.clip(c1 + scale * 1.98 * rng.standard_normal(n), -0.99, 0.99)
    out[3] = np.clip(c2 + scale * 1.98 * rng.standard_normal(n), -0.99, 0.99)
    out[4] = d + scale * np.diff(PRIOR_RANGES["dist"])[0] * rng.standard_normal(n)
    out[5] = phi + scale * 2 * np.pi * rng.standard_normal(n)
    out[6] = 0.0
    out[7] = np.arccos(np.clip(np.cos(inc) + scale * 2.0 * rng.standard_normal(n), -1, 1))
    out[8] = lam + scale * 2 * np.pi * rng.standard_normal(n)
    out[9] = np.arcsin(np.clip(np.sin(beta) + scale * 2.0 * rng.standard_normal(n), -1, 1))
    out[10] = psi + scale * np.pi * rng.standard_normal(n)
    out[11] = tref + scale * 0.8 * YRSID_SI * rng.standard_normal(n)
    return out
