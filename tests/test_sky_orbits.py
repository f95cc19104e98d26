"""CPU tests of the sky-frame utilities and of orbit tables given as arrays.

The frame conversions (utils/transform.py:136-310 of the reference) are pinned
by geometry stated independently of their formulas: the LISA-frame pole seen
from the SSB frame is n(alpha) = (-sin zeta cos alpha, -sin zeta sin alpha,
cos zeta), zeta = 60 degrees, alpha = Omega t, so sin(beta_L) = k . n(alpha);
the polarisation angle changes by the angle between the two frames' north
directions projected on the sky at the source; SSB -> LISA -> SSB returns the
input. The sky modes (:313-426) are pinned by their definitions.
"""

import numpy as np
import pytest

from bbhx_amd.utils import transform as T
from bbhx_amd.utils.constants import YRSID_SI
from bbhx_amd.utils.orbits import EqualArmlengthOrbits, TabulatedOrbits


def _sky(n, seed):
    rng = np.random.default_rng(seed)
    t = rng.uniform(0, 3 * YRSID_SI, n)
    lam = rng.uniform(0, 2 * np.pi, n)
    beta = np.arcsin(rng.uniform(-0.999, 0.999, n))
    psi = rng.uniform(0, np.pi, n)
    return t, lam, beta, psi


def _k(lam, beta):
    return np.stack([np.cos(beta) * np.cos(lam), np.cos(beta) * np.sin(lam), np.sin(beta)])


def test_lisa_latitude_is_the_height_over_the_tilted_pole():
    t, lam, beta, psi = _sky(2000, 1)
    _, lamL, betaL, _ = T.SSB_to_LISA(t, lam, beta, psi)
    a = T.ConstOmega * t
    z = np.pi / 3
    pole = np.stack([-np.sin(z) * np.cos(a), -np.sin(z) * np.sin(a), np.full_like(a, np.cos(z))])
    np.testing.assert_allclose(np.sin(betaL), (_k(lam, beta) * pole).sum(0), atol=1e-14)
    # the LISA-frame longitude is measured from the direction alpha (the Sun-LISA axis): a source in the
    # ecliptic plane at lam = alpha sits at lamL = alpha on the tilted equator's meridian
    _, lam0, beta0, _ = T.SSB_to_LISA(t, a, np.zeros_like(a), psi)
    np.testing.assert_allclose(np.cos(lam0 - a), 1.0, atol=1e-12)
    np.testing.assert_allclose(beta0, -z, atol=1e-12)


def test_polarisation_turns_by_the_angle_between_the_frames_norths():
    t, lam, beta, psi = _sky(2000, 2)
    _, lamL, betaL, psiL = T.SSB_to_LISA(t, lam, beta, psi)
    k = _k(lam, beta)
    a = T.ConstOmega * t
    z = np.pi / 3
    poles = (np.array([0.0, 0.0, 1.0])[:, None] * np.ones_like(a),
             np.stack([-np.sin(z) * np.cos(a), -np.sin(z) * np.sin(a), np.full_like(a, np.cos(z))]))
    norths = []
    for p in poles:
        v = p - (p * k).sum(0) * k
        norths.append(v / np.sqrt((v * v).sum(0)))
    cos_t = (norths[0] * norths[1]).sum(0)
    # polarisation angles are defined modulo pi: compare cos 2(psiL - psi) with cos 2 theta
    np.testing.assert_allclose(np.cos(2 * (psiL - psi)), 2 * cos_t**2 - 1, atol=1e-10)


def test_ssb_lisa_round_trip():
    t, lam, beta, psi = _sky(2000, 3)
    for t0 in (0.0, 0.7):
        tL, lamL, betaL, psiL = T.SSB_to_LISA(t, lam, beta, psi, t0=t0)
        t2, lam2, beta2, psi2 = T.LISA_to_SSB(tL, lamL, betaL, psiL, t0=t0)
        assert np.max(np.abs(t2 - t)) < 1e-4  # first-order light-time inverse vs second-order forward
        assert np.max(np.abs(np.angle(np.exp(1j * (lam2 - lam))))) < 1e-9
        assert np.max(np.abs(beta2 - beta)) < 1e-9
        assert np.max(np.abs(np.angle(np.exp(2j * (psi2 - psi))))) < 1e-9
        assert np.all((lam2 >= 0) & (lam2 < 2 * np.pi)) and np.all((psi2 >= 0) & (psi2 < np.pi))


def test_modpi_and_mod2pi():
    x = np.array([-7.0, -np.pi, 0.0, 1.0, np.pi, 10.0])
    assert np.all((T.modpi(x) >= 0) & (T.modpi(x) < np.pi))
    assert np.all((T.mod2pi(x) >= 0) & (T.mod2pi(x) < 2 * np.pi))
    np.testing.assert_allclose(np.cos(2 * T.modpi(x)), np.cos(2 * x), atol=1e-12)
    np.testing.assert_allclose(np.cos(T.mod2pi(x)), np.cos(x), atol=1e-12)


def _coords(n=3):
    rng = np.random.default_rng(4)
    c = rng.uniform(0.1, 1.0, (n, 12))
    c[:, 7] = rng.uniform(0, np.pi, n)        # inc
    c[:, 8] = rng.uniform(0, 2 * np.pi, n)    # lam
    c[:, 9] = rng.uniform(-1.2, 1.2, n)       # beta
    c[:, 10] = rng.uniform(0, np.pi, n)       # psi
    return c


def test_sky_modes_lat():
    c = _coords()
    out = T.mbh_sky_mode_transform(c, kind="lat")
    assert out.shape == (6, 12)
    for r in range(6):
        src = c[r % 3]
        if r % 2 == 0:
            np.testing.assert_array_equal(out[r], src)
        else:
            np.testing.assert_allclose(out[r, [7, 9, 10]], [np.pi - src[7], -src[9], np.pi - src[10]])
            np.testing.assert_array_equal(out[r, [0, 1, 2, 3, 4, 5, 6, 8, 11]], src[[0, 1, 2, 3, 4, 5, 6, 8, 11]])
    # twice in place is the identity; cos_i flips the sign of cos(inc)
    d = out.copy()
    T.mbh_sky_mode_transform(d, kind="lat", inplace=True)
    np.testing.assert_allclose(d, np.tile(c, (2, 1)))
    e = T.mbh_sky_mode_transform(np.cos(c), kind="lat", cos_i=True, ind_map=dict(inc=7, lam=8, beta=9, psi=10))
    np.testing.assert_allclose(e[1, 7], -np.cos(c[1, 7]))


def test_sky_modes_long_and_both():
    c = _coords(4)
    out = T.mbh_sky_mode_transform(c, kind="long")
    assert out.shape == (16, 12)
    for r in range(16):
        k, src = r % 4, c[r % 4]
        np.testing.assert_allclose(out[r, 8], (src[8] + k * np.pi / 2) % (2 * np.pi))
        np.testing.assert_allclose(out[r, 10], (src[10] + k * np.pi / 2) % np.pi)
    c = _coords(8)
    out = T.mbh_sky_mode_transform(c, kind="both")
    assert out.shape == (64, 12)
    for r in range(64):
        src = c[r % 8]
        k, flip = (r % 8) // 2, r % 2
        psi = np.pi - src[10] if flip else src[10]
        np.testing.assert_allclose(out[r, 10], (psi + k * np.pi / 2) % np.pi, atol=1e-15)
        np.testing.assert_allclose(out[r, 8], (src[8] + k * np.pi / 2) % (2 * np.pi) if k else src[8])
        np.testing.assert_allclose(out[r, 9], -src[9] if flip else src[9])


def test_sky_modes_argument_errors():
    c = _coords(3)
    with pytest.raises(ValueError):
        T.mbh_sky_mode_transform(c, kind="sideways")
    with pytest.raises(ValueError):
        T.mbh_sky_mode_transform(c, ind_map=[7, 8, 9, 10])
    with pytest.raises(ValueError):
        T.mbh_sky_mode_transform(c, kind="lat", inplace=True)


def test_tabulated_orbits_from_arrays_and_file(tmp_path):
    eq = EqualArmlengthOrbits(dt=1e5, t_max=1.0 * YRSID_SI)
    tab = TabulatedOrbits(**eq.to_arrays())
    for a, b in ((tab.x_base, eq.x_base), (tab.n_base, eq.n_base), (tab.ltt_base, eq.ltt_base),
                 (tab.t_base, eq.t_base)):
        np.testing.assert_array_equal(a, b)
    assert tab.as_host_dict()["N"] == eq.N and tab.armlength == eq.armlength
    path = tmp_path / "orbits.npz"
    tab.save(path)
    back = TabulatedOrbits.from_file(path)
    np.testing.assert_array_equal(back.x_base, eq.x_base)
    assert back.LINKS == eq.LINKS and back.dt == eq.dt
    # the response's argument check (fastfdresponse.py:107-134) accepts its arguments
    args = back.pycppdetector_args_on("cpu")
    assert len(args) == 9 and isinstance(args[0], float) and isinstance(args[1], int)


@pytest.mark.parametrize("change", ["dt", "shape", "links", "sc", "norm", "nan"])
def test_tabulated_orbits_reject_bad_tables(change):
    kw = EqualArmlengthOrbits(dt=1e5, t_max=0.2 * YRSID_SI).to_arrays()
    if change == "dt":
        kw["dt"] = 0.0
    elif change == "shape":
        kw["n_arr"] = kw["n_arr"][:, :5]
    elif change == "links":
        kw["links"] = np.array([12, 23, 31, 13, 32, 12])
    elif change == "sc":
        kw["sc_r"] = np.array([2, 2, 3, 1, 3, 2])
    elif change == "norm":
        kw["n_arr"] = kw["n_arr"] * 1.01
    else:
        kw["x_arr"] = kw["x_arr"].copy()
        kw["x_arr"][3, 1, 2] = np.nan
    with pytest.raises(ValueError):
        TabulatedOrbits(**kw)
