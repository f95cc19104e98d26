"""Pins of the oracle's LISA response and observation-window refit (CPU only).

The response oracle (oracle/oracle_response.cpp) restates Response.cu; these
tests check it against constructions that share none of its code:

* the TDI transfer functions against a time-domain evaluation of the
  single-link Doppler observables of a monochromatic plane wave on a frozen,
  exactly equilateral constellation, carried through the Michelson delays of
  TDI in extended precision (``np.clongdouble``). Its only inputs are the
  defining formula of the single-link response,
      y_sr(t) = n^T [h(t - L/c, p_s) - h(t, p_r)] n / (2 (1 - k.n)),
      h(t, x) = H exp(-2 pi i f (t - k.x / c)),  n = (p_r - p_s) / L,
  the first-generation Michelson combination with its (1 - D^2) and, for
  TDI 2, (1 - D^4) factors (D a delay by L/c; the reference applies the TDI-2
  factor to A, E, T only and returns first-generation X, Y, Z whatever tdi2
  says, Response.cu:209-224, :247-258), and the link labels of
  Response.cu:103-128 (the orbit row of link 12 holds n1, the unit vector of
  the arm opposite spacecraft 1 oriented cyclically, n1 = (p3 - p2)/L,
  n2 = (p1 - p3)/L, n3 = (p2 - p1)/L);
* the spin-weight -2 harmonics against the Wigner small-d closed form
  (sqrt((2l+1)/4pi) d^l_{m,2}(theta) e^{i m phi}) and against unit norm and
  orthogonality on the sphere by quadrature;
* the polarisation tensors against a geometric construction in the sky basis
  (e_lambda, e_beta) and against transversality, tracelessness, handedness and
  their spin-2 behaviour under psi -> psi + pi/4, psi + pi/2;
* the window refit (phenomhm.py:509-614) against the root of the oracle's own
  t(f) found by bisection, and the refitted grid against logspace.
"""

import math

import numpy as np
import pytest

import oracle
from oracle.pipeline import OracleWaveform
from tests.params import MODES, YRSID_SI, draw

C_SI = 299792458.0
AU_SI = 149597870700.0
ARM = 2.5e9
LINKS = [12, 23, 31, 13, 32, 21]


# ------------------------------------------------------------------ harmonics


def wigner_small_d(l, mp, m, beta):
    """d^l_{mp m}(beta) by the explicit Wigner sum."""
    f = math.factorial
    out = 0.0
    for k in range(0, 2 * l + 1):
        a, b, c, d = l + m - k, k, l - k - mp, k - m + mp
        if min(a, b, c, d) < 0:
            continue
        num = (-1) ** (k - m + mp) * math.sqrt(f(l + mp) * f(l - mp) * f(l + m) * f(l - m))
        out += num / (f(a) * f(b) * f(c) * f(d)) * np.cos(beta / 2) ** (2 * l - 2 * k + m - mp) * np.sin(
            beta / 2) ** (2 * k - m + mp)
    return out


def sylm_closed(l, m, theta, phi):
    """-2Y_lm from the Wigner small-d: sqrt((2l+1)/4pi) d^l_{m,2}(theta) e^{i m phi}."""
    return np.sqrt((2 * l + 1) / (4 * np.pi)) * wigner_small_d(l, m, 2, theta) * np.exp(1j * m * phi)


LM_ALL = [(l, s * m) for (l, m) in MODES for s in (1, -1)]


def test_sylm_matches_wigner_closed_form():
    rng = np.random.default_rng(11)
    for l, m in LM_ALL:
        for theta, phi in zip(rng.uniform(0, np.pi, 7), rng.uniform(0, 2 * np.pi, 7)):
            got = oracle.sylm(l, m, theta, phi)
            want = sylm_closed(l, m, theta, phi)
            assert abs(got - want) <= 1e-14 * max(1.0, abs(want)), (l, m, theta, got, want)


def test_sylm_orthonormal_on_the_sphere():
    """Gauss-Legendre in cos(theta) x uniform in phi is exact for these
    band-limited products; the Gram matrix of the 12 harmonics is the identity."""
    x, w = np.polynomial.legendre.leggauss(40)
    th = np.arccos(x)
    ph = np.arange(32) * (2 * np.pi / 32)
    Y = np.array([[[oracle.sylm(l, m, t, p) for p in ph] for t in th] for (l, m) in LM_ALL])
    wq = w[:, None] * (2 * np.pi / 32)
    gram = np.einsum("ats,bts,ts->ab", Y, Y.conj(), np.broadcast_to(wq, Y.shape[1:]))
    np.testing.assert_allclose(gram, np.eye(len(LM_ALL)), atol=1e-13)


# ------------------------------------------------------ polarisation tensors


def sky_basis(lam, beta, psi):
    """u = -cos(psi) e_lam + sin(psi) e_beta, v = cos(psi) e_beta + sin(psi) e_lam,
    k = -(cos beta cos lam, cos beta sin lam, sin beta), with e_lam, e_beta the
    unit vectors of increasing lambda and beta at the source's sky position
    (the frame of Response.cu:540-567, built here from the sky geometry)."""
    e_lam = np.array([-np.sin(lam), np.cos(lam), 0.0])
    e_beta = np.array([-np.sin(beta) * np.cos(lam), -np.sin(beta) * np.sin(lam), np.cos(beta)])
    k = -np.array([np.cos(beta) * np.cos(lam), np.cos(beta) * np.sin(lam), np.sin(beta)])
    u = -np.cos(psi) * e_lam + np.sin(psi) * e_beta
    v = np.cos(psi) * e_beta + np.sin(psi) * e_lam
    return u, v, k


def test_polarization_tensors_geometry():
    rng = np.random.default_rng(12)
    for lam, beta, psi in zip(rng.uniform(0, 2 * np.pi, 9), np.arcsin(rng.uniform(-1, 1, 9)),
                              rng.uniform(0, np.pi, 9)):
        k, hp, hc = oracle.polarization(lam, beta, psi)
        u, v, k2 = sky_basis(lam, beta, psi)
        np.testing.assert_allclose(k, k2, atol=1e-15)
        np.testing.assert_allclose(hp, np.outer(u, u) - np.outer(v, v), atol=1e-15)
        np.testing.assert_allclose(hc, np.outer(u, v) + np.outer(v, u), atol=1e-15)
        np.testing.assert_allclose(np.cross(u, v), k, atol=1e-15)  # right-handed (u, v, k)
        for H in (hp, hc):
            np.testing.assert_allclose(H, H.T, atol=1e-15)
            assert abs(np.trace(H)) < 1e-15
            np.testing.assert_allclose(H @ k, 0.0, atol=1e-15)
        np.testing.assert_allclose(np.sum(hp * hp), 2.0, rtol=1e-14)
        np.testing.assert_allclose(np.sum(hp * hc), 0.0, atol=1e-15)
        # spin 2: a quarter turn of psi exchanges the polarisations, half a turn flips both
        _, hp4, hc4 = oracle.polarization(lam, beta, psi + np.pi / 4)
        _, hp2, hc2 = oracle.polarization(lam, beta, psi + np.pi / 2)
        np.testing.assert_allclose(hp4, hc, atol=1e-14)
        np.testing.assert_allclose(hc4, -hp, atol=1e-14)
        np.testing.assert_allclose(hp2, -hp, atol=1e-14)
        np.testing.assert_allclose(hc2, -hc, atol=1e-14)


# ------------------------------------------- TDI from the time-domain definition


def frozen_constellation(seed):
    """Spacecraft positions of an exactly equilateral triangle of side ARM about
    a guiding centre at 1 AU, in a plane tilted by 60 degrees, and an orbit
    table (N = 2 rows, constant in time) carrying them and the link vectors in
    the labelling of Response.cu:103-128."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0, 2 * np.pi)
    centre = AU_SI * np.array([np.cos(a), np.sin(a), 0.0])
    radial = np.array([np.cos(a), np.sin(a), 0.0])
    z = np.array([0.0, 0.0, 1.0])
    e1 = np.cos(np.pi / 3) * radial + np.sin(np.pi / 3) * z
    e2 = np.cross(z, radial)
    th0 = rng.uniform(0, 2 * np.pi)
    r = ARM / np.sqrt(3.0)
    p = np.array([centre + r * (np.cos(th0 + 2 * np.pi * i / 3) * e1 + np.sin(th0 + 2 * np.pi * i / 3) * e2)
                  for i in range(3)])
    n1, n2, n3 = (p[2] - p[1]) / ARM, (p[0] - p[2]) / ARM, (p[1] - p[0]) / ARM
    rows = {12: n1, 23: n2, 31: n3, 13: -n1, 32: -n2, 21: -n3}
    n_arr = np.array([[rows[lk] for lk in LINKS]] * 2)
    x_arr = np.array([p, p])
    orbit = dict(dt=1e4, N=2, n_arr=n_arr, x_arr=x_arr, links=np.asarray(LINKS, np.int32), armlength=ARM)
    return p, orbit


def tdi_time_domain(f, H, k, p, tdi2, tag):
    """Transfer functions of the three channels: each observable evaluated as a
    function of t from the single-link definition, with explicit delays, then
    divided by exp(-2 pi i f t). Extended precision throughout."""
    ld = np.longdouble
    f = ld(f)
    c = ld(C_SI)
    L = ld(ARM)
    two_pi_i = ld(2) * ld(np.pi) * 1j
    Hq = H.astype(np.clongdouble)
    kq = k.astype(ld)
    P = p.astype(ld)

    def h_proj(n, t, x):
        return (n @ Hq @ n) * np.exp(-two_pi_i * f * (t - (kq @ x) / c))

    def y(s, r, t):  # link emitted by spacecraft s, received by r
        n = (P[r - 1] - P[s - 1]) / L
        return (h_proj(n, t - L / c, P[s - 1]) - h_proj(n, t, P[r - 1])) / (2 * (1 - kq @ n))

    def michelson(i, j, k_, t):  # first generation, equal arms, vertex i, arms to j and k_
        return (y(j, i, t) + y(i, j, t - L / c)) - (y(k_, i, t) + y(i, k_, t - L / c))

    def full(i, j, k_, t):
        x = michelson(i, j, k_, t - 2 * L / c) - michelson(i, j, k_, t)
        if tdi2 and tag == 2:
            x = x - (michelson(i, j, k_, t - 6 * L / c) - michelson(i, j, k_, t - 4 * L / c))
        return x

    t0 = ld(12.5)
    X, Y, Z = (full(1, 2, 3, t0), full(2, 3, 1, t0), full(3, 1, 2, t0))
    if tag == 2:
        X, Y, Z = (Z - X) / np.sqrt(ld(2)), (X - 2 * Y + Z) / np.sqrt(ld(6)), (X + Y + Z) / np.sqrt(ld(3))
    carrier = np.exp(-two_pi_i * f * t0)
    return np.array([X / carrier, Y / carrier, Z / carrier], dtype=np.complex128)


@pytest.mark.parametrize("tag,tdi2", [(1, False), (2, False), (1, True), (2, True)])
def test_response_matches_time_domain_tdi(tag, tdi2):
    p, orbit = frozen_constellation(21 + tag + 2 * tdi2)
    B, L = 3, 40
    prm = draw(B, 13)
    phi_ref, inc, lam, beta, psi = prm[5], prm[7], prm[8], prm[9], prm[10]
    M = len(MODES)
    ells = np.array([l for l, m in MODES], np.int32)
    mms = np.array([m for l, m in MODES], np.int32)
    # 1e-3 .. 0.3 Hz: pi f L / c from 0.026 to 7.9, across the transfer function's zeros
    fgrid = np.logspace(-3, np.log10(0.3), L)
    freqs = np.tile(fgrid, (B, M, 1))
    nper = B * M * L
    buf = np.zeros(9 * nper)
    buf[:nper] = 1.0
    buf[2 * nper:3 * nper] = 3.0e3  # t(f): any time, the table is constant
    oracle.lisa_response(buf, ells, mms, freqs.ravel(), phi_ref, inc, lam, beta, psi, tag, False, tdi2, M, L, B, 1,
                         orbit)
    b = buf.reshape(9, B, M, L)
    got = np.array([b[3] + 1j * b[4], b[5] + 1j * b[6], b[7] + 1j * b[8]])  # (3, B, M, L)
    for bi in range(B):
        u, v, k = sky_basis(lam[bi], beta[bi], psi[bi])
        ep, ec = np.outer(u, u) - np.outer(v, v), np.outer(u, v) + np.outer(v, u)
        for mi, (l, m) in enumerate(MODES):
            ylm = sylm_closed(l, m, inc[bi], phi_ref[bi])
            yl_m = (-1) ** l * np.conj(sylm_closed(l, -m, inc[bi], phi_ref[bi]))
            H = 0.5 * (ylm + yl_m) * ep + 0.5j * (ylm - yl_m) * ec
            want = np.array([tdi_time_domain(f, H, k, p, tdi2, tag) for f in fgrid]).T  # (3, L)
            for ch in range(3):
                scale = np.max(np.abs(want[ch]))
                np.testing.assert_allclose(got[ch, bi, mi], want[ch], rtol=1e-9, atol=1e-10 * scale,
                                           err_msg=f"tag {tag} tdi2 {tdi2} binary {bi} mode {(l, m)} channel {ch}")


# ------------------------------------------------------------- window refit


def _tf_of(w, args, f, mode_index):
    """The oracle's t(f) of one mode at frequencies f (direct evaluation)."""
    m1, m2, c1, c2, d, phi, fref, tref = args
    w.run_wave(m1, m2, c1, c2, d, phi, fref, tref, len(f), freqs=np.tile(f, (1, len(MODES), 1)), modes=MODES)
    n = len(MODES) * len(f)
    return w.buffer[2 * n:3 * n].reshape(len(MODES), -1)[mode_index]


def test_window_refit_brackets_the_root_and_grid_is_logspace():
    """phenomhm.py:509-614: where t(f) crosses 0 (and Tobs) the grid end is
    moved to the linear interpolation of f between the bracketing grid points;
    the refitted grid is logspace between the new ends."""
    from bbhx_amd.utils.orbits import EqualArmlengthOrbits

    orb = EqualArmlengthOrbits(dt=1e4, t_max=3 * YRSID_SI).as_host_dict()
    w = OracleWaveform(orb)
    # t_ref early in the year: every mode starts before t = 0
    prm = draw(12, 31, mtot=(1e5, 1e6), tref=(0.02, 0.05))
    L = 256
    checked_lo = checked_hi = 0
    for bi in range(prm.shape[1]):
        Tobs = 0.6 * prm[11, bi]  # before merger: every mode crosses it
        args = (prm[0, bi:bi + 1], prm[1, bi:bi + 1], prm[2, bi:bi + 1], prm[3, bi:bi + 1], prm[4, bi:bi + 1],
                np.zeros(1), prm[6, bi:bi + 1], prm[11, bi:bi + 1])
        # the first pass on the default grid, as amp_phase runs it
        w.run_wave(*args, L, modes=MODES)
        f0 = w.freqs.reshape(len(MODES), L).copy()
        n = len(MODES) * L
        tf0 = w.buffer[2 * n:3 * n].reshape(len(MODES), L).copy()
        w.amp_phase(*args, L, modes=MODES, Tobs=np.array([Tobs]))
        fn = w.freqs.reshape(len(MODES), L).copy()
        for mi in range(len(MODES)):
            np.testing.assert_allclose(fn[mi], np.logspace(np.log10(fn[mi, 0]), np.log10(fn[mi, -1]), L), rtol=1e-15)
            for target, end in ((0.0, 0), (Tobs, -1)):
                cross = np.nonzero((tf0[mi, :-1] < target) & (tf0[mi, 1:] > target))[0]
                if end == 0 and tf0[mi].min() >= 0.0 or end == -1 and tf0[mi].max() <= Tobs:
                    # untouched where no crossing (up to logspace's 10**log10 round trip)
                    assert abs(fn[mi, end] - f0[mi, end]) <= 1e-14 * f0[mi, end]
                    continue
                assert cross.size == 1
                s = cross[0]
                lo, hi = f0[mi, s], f0[mi, s + 1]
                assert lo <= fn[mi, end] <= hi
                # root of the oracle's own t(f) by bisection inside the bracket
                a, bnd = lo, hi
                for _ in range(45):
                    mid = 0.5 * (a + bnd)
                    if _tf_of(w, args, np.array([mid]), mi)[0] < target:
                        a = mid
                    else:
                        bnd = mid
                root = 0.5 * (a + bnd)
                # linear interpolation of a smooth monotone t(f): the root error
                # is second order in the bracket width
                assert abs(fn[mi, end] - root) < 0.05 * (hi - lo), (bi, mi, end, fn[mi, end], root, lo, hi)
                if end == 0:
                    checked_lo += 1
                else:
                    checked_hi += 1
    assert checked_lo >= 60 and checked_hi >= 60, (checked_lo, checked_hi)
