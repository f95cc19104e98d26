"""FP32 rounding budget of the heterodyned sums (TEST INFRASTRUCTURE, CPU only).

Question answered: can the heterodyned likelihood's sums run in FP32 (products
and compensated accumulation in FP32, everything before them in FP64) and stay
within |dlogL| < 1e-3 of FP64 on the heterodyned benchmark batch (one reference
drawn from the small-batch distribution, templates offset by N(0, 1e-3 range),
K = 128 log-spaced nodes, 2^20 dense bins over one year, six modes, TDI-2 AET)?

The residuals r = h / h0 at the nodes and the constants A0, A1, B0, B1 come
from the FP64 oracle (likelihood.py:499-533 for the constants,
Likelihood.cu:52-111 for the segment sums). Two FP32 forms of the sums are
evaluated, with FP32 products and a Neumaier-compensated FP32 running sum per
template (the order of the reference's loop: channel, then segment):

* plain: the reference's terms A0 conj(r0) + A1 conj(r1) and
  B0 |r0|^2 + 2 B1 Re(r0 conj(r1)), with r0, r1 and the constants rounded to FP32;
* delta: sum A0 and sum B0 kept in FP64 (the reference template's own d_h, h_h),
  and only the terms in dr0 = r0 - 1 and r1 accumulated in FP32.

r1 and A1, B1 are rescaled by the segment width so that every FP32 quantity is
O(1) relative to its neighbours (no change in exact arithmetic).

    python oracle/precision_budget.py [templates]
"""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))

from bbhx_amd.utils.constants import YRSID_SI  # noqa: E402
from bbhx_amd.utils.orbits import EqualArmlengthOrbits  # noqa: E402
from bbhx_amd.utils.sensitivity import AET1SensitivityMatrix  # noqa: E402
from oracle.pipeline import OracleHeterodyne, OracleWaveform  # noqa: E402
from tests.params import MODES, draw, perturb  # noqa: E402


def neumaier32(terms):
    """Neumaier-compensated sum in FP32 over axis 1 of (B, n) float32 terms,
    returned as s + c in FP64 (an FP32 result alone holds 3e5 only to 0.02)."""
    s = np.zeros(terms.shape[0], np.float32)
    c = np.zeros(terms.shape[0], np.float32)
    for k in range(terms.shape[1]):
        x = terms[:, k]
        t = (s + x).astype(np.float32)
        big = np.abs(s) >= np.abs(x)
        c = (c + np.where(big, (s - t) + x, (x - t) + s)).astype(np.float32)
        s = t
    return s.astype(np.float64) + c.astype(np.float64)


def csum32(z):
    return neumaier32(z.real.astype(np.float32)) + 1j * neumaier32(z.imag.astype(np.float32))


def main(B=512, K=128, seed=1000):
    ref = draw(16, seed)[:, 0]
    tmpl = perturb(ref, 4096, seed + 1, scale=1e-3)[:, :B]
    f = np.arange(1, 2**20 + 1) / YRSID_SI
    orbits = EqualArmlengthOrbits(dt=1e4, t_max=3.0 * YRSID_SI)
    owave = OracleWaveform(orbits.as_host_dict(), tdi2=True)
    sens = AET1SensitivityMatrix(f)
    psd = np.array([sens[0], sens[1], sens[2]])
    inj = owave(*ref, freqs=f, length=1024, modes=MODES, direct=False, fill=True, squeeze=True)[0]
    rng = np.random.default_rng(seed + 1)
    df = f[1] - f[0]
    data = inj + np.sqrt(psd / (4 * df)) * (rng.standard_normal(psd.shape) + 1j * rng.standard_normal(psd.shape))
    psd_fn = lambda x: [AET1SensitivityMatrix(x)[i] for i in range(3)]  # noqa: E731
    het = OracleHeterodyne(owave, f, data, ref, K, psd_fn, reference_length=8096, modes=MODES)
    _, dh_o, hh_o = het.get_ll(tmpl)
    h = het.wave(*tmpl, freqs=het.freqs, direct=True, compress=True, squeeze=False, **het.wkw)
    r = h / het.h0_sparse                                   # (B, 3, K)
    A0, A1, B0, B1 = het.data_constants.reshape(4, 3, K)[:, :, 1:]
    x = het.freqs
    w = x[1:] - x[:-1]
    r1w = r[:, :, 1:] - r[:, :, :-1]                        # r1 * w
    r0 = r[:, :, 1:] - r1w / 2                              # r at the segment midpoint
    # FP64 sums in the same form (checked against the oracle's hdyn)
    dh64 = (A0 * r0.conj() + (A1 / w) * r1w.conj()).sum((1, 2))
    hh64 = (B0.real * np.abs(r0) ** 2 + 2 * (B1.real / w) * (r0 * r1w.conj()).real).sum((1, 2))
    assert np.max(np.abs(dh64 - dh_o)) <= 1e-12 * np.max(np.abs(dh_o))
    assert np.max(np.abs(hh64 - hh_o)) <= 1e-12 * np.max(np.abs(hh_o))
    ll = lambda dh, hh: -0.5 * (het.reference_d_d + hh - 2 * dh.real)  # noqa: E731
    ll64 = ll(dh64, hh64)

    c64, f32 = np.complex64, np.float32
    a0, a1, b0, b1 = A0.astype(c64), (A1 / w).astype(c64), B0.real.astype(f32), (B1.real / w).astype(f32)
    q1 = r1w.astype(c64)
    flat = lambda z: z.reshape(B, -1)  # noqa: E731  (channel-major, then segment)
    out = {}
    # plain FP32 terms
    q0 = r0.astype(c64)
    t_dh = a0 * q0.conj() + a1 * q1.conj()
    t_hh = b0 * (q0.real * q0.real + q0.imag * q0.imag) + f32(2) * b1 * (q0 * q1.conj()).real
    out["plain"] = ll(csum32(flat(t_dh)), neumaier32(flat(t_hh)))
    # delta form: sum A0, sum B0 in FP64
    d0 = (r0 - 1.0).astype(c64)
    t_dh = a0 * d0.conj() + a1 * q1.conj()
    t_hh = b0 * (f32(2) * d0.real + d0.real * d0.real + d0.imag * d0.imag) + f32(2) * b1 * (
        (f32(1) + d0) * q1.conj()).real
    dh = A0.sum() + csum32(flat(t_dh))
    hh = B0.real.sum() + neumaier32(flat(t_hh))
    out["delta"] = ll(dh, hh)

    print(f"{B} templates, K = {K}: h_h {hh64.min():.3g}-{hh64.max():.3g}, |d_h| up to {np.abs(dh64).max():.3g}, "
          f"|r0 - 1| median {np.median(np.abs(r0 - 1)):.3g}, max {np.abs(r0 - 1).max():.3g}")
    for name, v in out.items():
        d = np.abs(v - ll64)
        print(f"FP32 sums, {name:5s}: |dlogL| vs FP64 max {d.max():.3e}, median {np.median(d):.3e}, "
              f"{int(np.sum(d < 1e-3))} of {B} within 1e-3; max |dlogL| / h_h {np.max(d / hh64):.2e}")
    return out, ll64


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 512)
