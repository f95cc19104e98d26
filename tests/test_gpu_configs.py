"""The benchmark configurations at their own workloads, against the FP64 CPU
oracle.

* Direct likelihood, 16 templates (the first seeded draws), default grids of
  L = 1024 refitted to the observation window, 2^16 dense bins at df = 1/yr,
  injection of the first draw + Gaussian noise: kernel-by-kernel and fused HIP
  logL against the oracle on all 16.
* Uneven batching: the 8192-template ragged batch (total mass 1e4-1e8,
  sparse lengths 256-4096 from their observable bands) on 1 yr at 10 s
  (1,577,907 bins), fused: the oracle on 32 templates, among them the
  shortest and the longest grids; on all 8192, finite values, the
  Cauchy-Schwarz bound |d_h| <= sqrt(d_d h_h) on each template's band, and the
  FP32-bulk form within |dlogL| < 1e-3 of the FP64 form.
* Gradient, 1024 templates around the reference, K = 256, 2^20 bins: the
  adjoint gradient against fourth-order central differences of the oracle's
  heterodyned logL on 8 templates (1e-5 relative, plus the differences' own
  error estimated from two step sizes), finite everywhere, and the logL the
  gradient entry point returns equal to the forward's on all 1024.

logL bounds: |dlogL| <= 1e-6 + 1e-12 |logL| (at |logL| of 1e5-1e7 the FP64
rounding of the PhenomHM phases alone moves logL by ~1e-6: the oracle built
with FMA contraction differs from itself by that much).
"""

import numpy as np
import pytest

from tests.params import MODES, YRSID_SI, draw

pytestmark = pytest.mark.gpu


def _bound(ll_ref):
    return 1e-6 + 1e-12 * np.abs(ll_ref)


def test_direct_16_templates_1024_points_2e16_bins():
    from bbhx_amd.likelihood import Likelihood
    from bbhx_amd.utils.orbits import EqualArmlengthOrbits
    from bbhx_amd.utils.sensitivity import AET1SensitivityMatrix
    from bbhx_amd.waveformbuild import BBHWaveformFD
    from oracle.pipeline import OracleLikelihood, OracleWaveform

    p = draw(16, 1000)
    f = np.arange(1, 2**16 + 1) / YRSID_SI
    orbits = EqualArmlengthOrbits(dt=1e4, t_max=3.0 * YRSID_SI)
    owave = OracleWaveform(orbits.as_host_dict(), tdi2=True)
    sens = AET1SensitivityMatrix(f)
    psd = np.array([sens[0], sens[1], sens[2]])
    inj = owave(*p[:, 0], freqs=f, length=1024, modes=MODES, direct=False, fill=True, squeeze=True)[0]
    rng = np.random.default_rng(1001)
    df = f[1] - f[0]
    data = inj + np.sqrt(psd / (4 * df)) * (rng.standard_normal(psd.shape) + 1j * rng.standard_normal(psd.shape))
    wave = BBHWaveformFD(response_kwargs=dict(orbits=orbits, tdi2=True))
    like = Likelihood(wave, f, data, psd)
    ll_plain = like.get_ll(p, length=1024, modes=MODES)
    ll_fused = like.get_ll_fused(p, np.full(16, 1024, dtype=np.int32), modes=MODES)
    olike = OracleLikelihood(owave, f, data, psd)
    ll_o = np.array([olike.get_ll(p[:, b:b + 1], length=1024, modes=MODES)[0][0] for b in range(16)])
    for name, ll in (("kernel-by-kernel", ll_plain), ("fused", ll_fused)):
        d = np.abs(ll - ll_o)
        print(f"{name}: max |dlogL| {d.max():.3e} at |logL| {np.abs(ll_o[np.argmax(d)]):.3e}", flush=True)
        assert np.all(np.isfinite(ll))
        assert np.all(d <= _bound(ll_o)), name


def test_uneven_batch_8192_templates():
    import torch

    import bench
    from bbhx_amd.likelihood import Likelihood
    from bbhx_amd.utils.orbits import EqualArmlengthOrbits
    from bbhx_amd.utils.sensitivity import AET1SensitivityMatrix
    from bbhx_amd.waveformbuild import BBHWaveformFD
    from oracle.pipeline import OracleLikelihood, OracleWaveform

    n = int(YRSID_SI // 10.0)
    f = np.fft.rfftfreq(n, 10.0)[1:]
    assert f.size == 1577907
    p = draw(8192, 1002, mtot=(1e4, 1e8))
    lengths = bench.ragged_lengths(p, YRSID_SI)
    orbits = EqualArmlengthOrbits(dt=1e4, t_max=3.0 * YRSID_SI)
    wave = BBHWaveformFD(response_kwargs=dict(orbits=orbits, tdi2=True))
    sens = AET1SensitivityMatrix(f)
    psd = np.array([sens[0], sens[1], sens[2]])
    inj = wave(*p[:, 0], freqs=f, length=1024, modes=MODES, direct=False, fill=True)[0].cpu().numpy()
    rng = np.random.default_rng(1003)
    df = f[1] - f[0]
    data = inj + np.sqrt(psd / (4 * df)) * (rng.standard_normal(psd.shape) + 1j * rng.standard_normal(psd.shape))
    like = Likelihood(wave, f, data, psd)
    print(f"uneven batch: lengths {lengths.min()}-{lengths.max()}, {lengths.sum()} knots", flush=True)
    ll = like.get_ll_fused(p, lengths, modes=MODES)
    d_h, h_h = like.d_h.cpu().numpy(), like.h_h.cpu().numpy()
    torch.cuda.synchronize()
    assert np.all(np.isfinite(ll)) and np.all(np.isfinite(d_h)) and np.all(np.isfinite(h_h))
    assert np.all(h_h.real >= 0.0) and np.all(np.abs(h_h.imag) <= 1e-9 * np.maximum(h_h.real, 1.0))
    # Cauchy-Schwarz on each band: the band's <d|d> is at most the full d_d
    assert np.all(np.abs(d_h) <= np.sqrt(like.d_d * h_h.real) * (1 + 1e-12) + 1e-9)
    # the FP32-bulk form of the same batch: within 1e-3 of the FP64 form on every template
    plan = like.direct_plan(lengths, MODES)
    ll64 = like.get_ll_fused(p, lengths, modes=MODES, plan=plan)
    ll32 = like.get_ll_fused(p, lengths, modes=MODES, plan=plan, precision=1)
    plan.set_precision(0)
    d32 = np.abs(ll32 - ll64)
    print(f"FP32 bulk vs FP64: max |dlogL| {d32.max():.3e}, median {np.median(d32):.3e}, "
          f"{int(np.sum(h_h.real > 2.0e4))} templates recomputed in FP64", flush=True)
    assert np.all(d32 < 1e-3)
    del plan
    # oracle on 32 templates: the shortest and longest grids and 30 seeded others
    order = np.argsort(lengths, kind="stable")
    pick = np.unique(np.concatenate([order[:2], order[-2:], np.random.default_rng(7).choice(8192, 28, False)]))
    owave = OracleWaveform(orbits.as_host_dict(), tdi2=True)
    olike = OracleLikelihood(owave, f, data, psd)
    ll_o = np.array([olike.get_ll(p[:, b:b + 1], length=int(lengths[b]), modes=MODES)[0][0] for b in pick])
    d = np.abs(ll[pick] - ll_o)
    print(f"oracle on {pick.size} templates (L {lengths[pick].min()}-{lengths[pick].max()}): max |dlogL| "
          f"{d.max():.3e} at |logL| {np.abs(ll_o[np.argmax(d)]):.3e}", flush=True)
    assert np.all(d <= _bound(ll_o))


def test_gradient_1024_templates_256_nodes_2e20_bins():
    import bench
    from bbhx_amd.likelihood import HeterodynedLikelihood
    from bbhx_amd.utils.sensitivity import AET1SensitivityMatrix
    from oracle.pipeline import OracleHeterodyne, OracleWaveform
    from tests.test_gpu_grad import REL_DIST_STEP, STEPS, to_sampling

    prob = bench.build_problem(1024, 256, 20, seed=1003)
    het = prob["het"]
    X = to_sampling(prob["templates"])
    rows, ts, te = het.fused_inputs(het.sampling_to_params(X), {})
    ll_g, grad = (v.cpu().numpy() for v in het.get_ll_grad_device(rows, ts, te))
    grad = grad.T
    ll_f = het.get_ll_device(rows, ts, te).cpu().numpy()
    assert np.all(np.isfinite(grad)) and np.all(np.isfinite(ll_g))
    np.testing.assert_array_equal(ll_g, ll_f)
    owave = OracleWaveform(prob["orbits"].as_host_dict(), tdi2=True)
    psd_fn = lambda x: [AET1SensitivityMatrix(x)[i] for i in range(3)]  # noqa: E731
    ohet = OracleHeterodyne(owave, prob["f"], prob["data"], prob["ref"], 256, psd_fn, reference_length=8096,
                            modes=MODES)
    sel = np.arange(0, 1024, 128)  # 8 templates
    Xs = X[sel]
    g = grad[sel]

    def oll(Y):
        return ohet.get_ll(HeterodynedLikelihood.sampling_to_params(Y))[0]

    def fd(k, scale):
        h = scale * (STEPS[k] if k != 4 else REL_DIST_STEP * Xs[:, 4])
        e = np.zeros_like(Xs)
        e[:, k] = h
        return (8 * (oll(Xs + e) - oll(Xs - e)) - (oll(Xs + 2 * e) - oll(Xs - 2 * e))) / (12 * h)

    # Two steps h, 2h: their difference estimates the finite difference's own
    # error (its truncation, and the oracle's FP64 rounding of logL ~ 1e-6 at
    # |logL| ~ 3e6 amplified by 1/h, which for the smallest components exceeds
    # 1e-5 of them). Bound: 1e-5 relative plus three times that estimate.
    worst = 0.0
    for k in range(11):
        f1, f2 = fd(k, 1.0), fd(k, 2.0)
        err = np.abs(f1 - f2)
        rel = np.abs(g[:, k] - f1) / np.maximum(np.abs(f1), 1e-300)
        ok = np.abs(g[:, k] - f1) <= 1e-5 * np.abs(f1) + 3.0 * err
        worst = max(worst, float(rel.max()))
        print(f"component {k}: max relative difference {rel.max():.2e}, finite-difference error estimate "
              f"{np.max(err / np.maximum(np.abs(f1), 1e-300)):.2e} (|g| {np.abs(f1).min():.3e}-"
              f"{np.abs(f1).max():.3e}); {int(np.sum(rel <= 1e-5))} of {rel.size} within 1e-5", flush=True)
        assert np.all(ok), k
    print(f"gradient vs oracle finite differences: worst {worst:.2e}", flush=True)
