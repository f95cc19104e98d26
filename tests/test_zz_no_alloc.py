"""The allocation-free heterodyned entry point (bbhx_hip_hdyn_fused_noalloc)
against bbhx_hip_hdyn_fused, both called through the C ABI bindings on the
same inputs: equal bit for bit on two calls in a row, the second with a
smaller batch (the fixed record array keeps the first call's records, which
the second must not read), and a batch above the capacity is refused with
status 5 before any device work."""

import numpy as np
import pytest

from tests.params import MODES, YRSID_SI, draw, perturb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def het_small():
    import torch

    from bbhx_amd.likelihood import HeterodynedLikelihood
    from bbhx_amd.utils.orbits import EqualArmlengthOrbits
    from bbhx_amd.utils.sensitivity import AET1SensitivityMatrix
    from bbhx_amd.waveformbuild import BBHWaveformFD

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    ref = draw(16, 1000)[:, 0]
    f = np.arange(1, 2**15 + 1) * 4 / YRSID_SI
    orbits = EqualArmlengthOrbits(dt=1e4, t_max=3.0 * YRSID_SI)
    wave = BBHWaveformFD(response_kwargs=dict(orbits=orbits, tdi2=True))
    sens = AET1SensitivityMatrix(f)
    psd = np.array([sens[0], sens[1], sens[2]])
    data = wave(*ref, freqs=f, length=1024, modes=MODES, direct=False, fill=True)[0].cpu().numpy()
    rng = np.random.default_rng(11)
    df = f[1] - f[0]
    data = data + np.sqrt(psd / (4 * df)) * (rng.standard_normal(psd.shape) + 1j * rng.standard_normal(psd.shape))
    het = HeterodynedLikelihood(wave, f, data, ref, 96, sens_mat=sens, template_gen_kwargs=dict(modes=MODES),
                                reference_gen_kwargs=dict(length=2048, modes=MODES))
    return het, ref


def _call(het, rows, ts, te, noalloc):
    import torch

    from bbhx_amd import cutils

    gen = het.template_gen
    resp = gen.response_gen
    B = rows.shape[1]
    d_h = torch.full((B,), np.nan, dtype=torch.complex128, device=rows.device)
    h_h = torch.full((B,), np.nan, dtype=torch.complex128, device=rows.device)
    ll = torch.full((B,), np.nan, dtype=torch.float64, device=rows.device)
    ells = np.asarray([l for l, m in MODES], np.int32)
    mms = np.asarray([m for l, m in MODES], np.int32)
    cutils.hdyn_fused(d_h, h_h, rows, ts, te, ells, mms, het.freqs, het._h0_flat, het.data_constants,
                      gen.amp_phase_gen.qnm_packed, gen.amp_phase_gen.dspin, gen.amp_phase_gen.initial_t_val,
                      resp.cpp_response, resp.TDItag_int, resp.rescaled, resp.tdi2, het.include_last_segment,
                      logL=ll, d_d=het.reference_d_d, noalloc=noalloc)
    torch.cuda.synchronize()
    return d_h.cpu().numpy(), h_h.cpu().numpy(), ll.cpu().numpy()


def test_noalloc_matches_allocating_entry_point_twice(het_small):
    het, ref = het_small
    p = perturb(ref, 300, 5, scale=1e-3)
    for B in (300, 37):
        print(f"no-alloc vs allocating entry point: {B} templates, K = 96", flush=True)
        rows, ts, te = het.fused_inputs(p[:, :B])
        a = _call(het, rows, ts, te, noalloc=False)
        b = _call(het, rows, ts, te, noalloc=True)
        for x, y in zip(a, b):
            assert np.all(np.isfinite(x))
            np.testing.assert_array_equal(x, y)


def test_noalloc_refuses_batches_above_capacity(het_small):
    import torch

    from bbhx_amd import cutils

    het, ref = het_small
    B = cutils.NOALLOC_MAX_RECORDS // len(MODES) + 1
    rows, ts, te = het.fused_inputs(np.repeat(ref[:, None], B, axis=1))
    with pytest.raises(cutils.HipCallError, match="status 5"):
        _call(het, rows, ts, te, noalloc=True)
    torch.cuda.synchronize()
