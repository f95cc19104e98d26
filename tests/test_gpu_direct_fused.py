"""Fused direct likelihood of a ragged batch (per-template sparse grid lengths)
against the FP64 CPU oracle run template by template at each template's own
length, against the kernel-by-kernel GPU path, and under hipGraph capture.

Tolerance: |delta logL| < 1e-6 absolute (FP64 on both sides). The FP32-bulk
form (amplitude and transfer-function cubics, phase factor and mode sum in
FP32; phase, t(f) and the inner products in FP64): |delta logL| < 1e-3 against
the oracle, and d_h, h_h within 1e-5 of |h_h| (relative).
"""

import numpy as np
import pytest

from oracle.pipeline import OracleLikelihood, OracleWaveform
from tests.params import MODES, YRSID_SI, draw, perturb

pytestmark = pytest.mark.gpu

LENGTHS = np.array([256, 300, 512, 1000, 777, 1024, 64, 2048], dtype=np.int32)


@pytest.fixture(scope="module")
def direct_problem():
    from bbhx_amd import get_backend
    from bbhx_amd.likelihood import Likelihood
    from bbhx_amd.utils.orbits import EqualArmlengthOrbits
    from bbhx_amd.utils.sensitivity import AET1SensitivityMatrix
    from bbhx_amd.waveformbuild import BBHWaveformFD

    backend = get_backend("hip")
    orbits = EqualArmlengthOrbits(dt=1e4, t_max=3.0 * YRSID_SI)
    ref = draw(1, 1001)[:, 0]
    ref[0], ref[1] = 1.3e6, 4.1e5
    ref[11] = 0.62 * YRSID_SI
    f = np.arange(1, 2**15 + 1) * (4.0 / YRSID_SI)
    owave = OracleWaveform(orbits.as_host_dict(), tdi2=True)
    inj = owave(*ref, freqs=f, length=1024, modes=MODES, direct=False, fill=True, squeeze=True)[0]
    sens = AET1SensitivityMatrix(f)
    psd = np.array([sens[0], sens[1], sens[2]])
    rng = np.random.default_rng(1001)
    df = f[1] - f[0]
    data = inj + np.sqrt(psd / (4 * df)) * (rng.standard_normal(psd.shape) + 1j * rng.standard_normal(psd.shape))
    wave = BBHWaveformFD(response_kwargs=dict(orbits=orbits, tdi2=True), force_backend=backend)
    like = Likelihood(wave, f, data, psd, force_backend=backend)
    olike = OracleLikelihood(owave, f, data, psd)
    p = perturb(ref, LENGTHS.size, 22, scale=1e-5)
    return dict(like=like, olike=olike, p=p)


def test_ragged_fused_matches_oracle_per_template(direct_problem):
    like, olike, p = direct_problem["like"], direct_problem["olike"], direct_problem["p"]
    ll = like.get_ll_fused(p, LENGTHS, modes=MODES)
    ll_o = np.array([olike.get_ll(p[:, b:b + 1], length=int(L), modes=MODES)[0][0] for b, L in enumerate(LENGTHS)])
    print("fused", ll, "oracle", ll_o, flush=True)
    assert np.all(np.isfinite(ll))
    assert np.max(np.abs(ll - ll_o)) < 1e-6


def test_ragged_fp32_bulk_matches_oracle(direct_problem):
    like, olike, p = direct_problem["like"], direct_problem["olike"], direct_problem["p"]
    plan = like.direct_plan(LENGTHS, MODES)
    ll64 = like.get_ll_fused(p, LENGTHS, modes=MODES, plan=plan, precision=0)
    dh64, hh64 = like.d_h.cpu().numpy(), like.h_h.cpu().numpy()
    ll32 = like.get_ll_fused(p, LENGTHS, modes=MODES, plan=plan, precision=1)
    dh32, hh32 = like.d_h.cpu().numpy(), like.h_h.cpu().numpy()
    ll_o = np.array([olike.get_ll(p[:, b:b + 1], length=int(L), modes=MODES)[0][0] for b, L in enumerate(LENGTHS)])
    err = np.abs(ll32 - ll_o)
    scale = np.maximum(np.abs(hh64.real), 1.0)
    print("fp32 bulk", ll32, "oracle", ll_o, "|dlogL|", err, "h_h", hh64.real, flush=True)
    assert np.all(np.isfinite(ll32))
    assert np.max(err) < 1e-3
    assert np.max(np.abs(dh32 - dh64) / scale) < 1e-5
    assert np.max(np.abs(hh32 - hh64) / scale) < 1e-5
    # back to FP64 on the same plan: identical to the first FP64 call
    ll64b = like.get_ll_fused(p, LENGTHS, modes=MODES, plan=plan, precision=0)
    assert np.array_equal(ll64, ll64b)


def test_ragged_fused_matches_plain_gpu_path(direct_problem):
    like, p = direct_problem["like"], direct_problem["p"]
    lens = np.full(4, 512, dtype=np.int32)
    ll_f = like.get_ll_fused(p[:, :4], lens, modes=MODES)
    ll_p = like.get_ll(p[:, :4], length=512, modes=MODES)
    assert np.max(np.abs(ll_f - ll_p)) < 1e-6, (ll_f, ll_p)


def test_ragged_fused_graph_replays(direct_problem):
    import torch

    from bbhx_amd.likelihood import fused_param_rows

    like, p = direct_problem["like"], direct_problem["p"]
    plan = like.direct_plan(LENGTHS, MODES)
    rows, ts, te = fused_param_rows(like.waveform_gen, p, 0.0, 1.0, like.device)
    flags = torch.zeros(rows.shape[1], dtype=torch.int32, device=like.device)
    s_rows, s_ts, s_te = rows.clone(), ts.clone(), te.clone()
    # warm-up on a side stream, then capture
    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(st):
        flags.zero_()
        like.get_ll_fused_device(plan, s_rows, s_ts, s_te, flags, MODES)
    torch.cuda.current_stream().wait_stream(st)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        flags.zero_()
        out = like.get_ll_fused_device(plan, s_rows, s_ts, s_te, flags, MODES)
    torch.cuda.synchronize()
    rng = np.random.default_rng(3)
    for rep in range(3):
        q = p.copy()
        q[11] += 30.0 * rng.standard_normal(q.shape[1])
        q[5] += 0.1 * rng.standard_normal(q.shape[1])
        r2, t2, e2 = fused_param_rows(like.waveform_gen, q, 0.0, 1.0, like.device)
        eager = like.get_ll_fused(q, LENGTHS, modes=MODES, plan=plan)  # eager work between replays
        torch.cuda.synchronize()
        s_rows.copy_(r2)
        s_ts.copy_(t2)
        s_te.copy_(e2)
        g.replay()
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        print("replay", rep, got, eager, flush=True)
        assert int(flags.sum().item()) == 0
        assert np.max(np.abs(got - eager)) < 1e-9


def test_ragged_fp32_bulk_graph_replays(direct_problem):
    """The FP32-bulk form captured in a hipGraph (records allocated before the
    capture): replays with new inputs match eager calls (1e-9, as the FP64
    form's graph test)."""
    import torch

    from bbhx_amd.likelihood import fused_param_rows

    like, p = direct_problem["like"], direct_problem["p"]
    plan = like.direct_plan(LENGTHS, MODES)
    plan.set_precision(1)
    rows, ts, te = fused_param_rows(like.waveform_gen, p, 0.0, 1.0, like.device)
    flags = torch.zeros(rows.shape[1], dtype=torch.int32, device=like.device)
    s_rows, s_ts, s_te = rows.clone(), ts.clone(), te.clone()
    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(st):
        flags.zero_()
        like.get_ll_fused_device(plan, s_rows, s_ts, s_te, flags, MODES)
    torch.cuda.current_stream().wait_stream(st)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        flags.zero_()
        out = like.get_ll_fused_device(plan, s_rows, s_ts, s_te, flags, MODES)
    torch.cuda.synchronize()
    rng = np.random.default_rng(4)
    for rep in range(3):
        q = p.copy()
        q[11] += 30.0 * rng.standard_normal(q.shape[1])
        q[5] += 0.1 * rng.standard_normal(q.shape[1])
        r2, t2, e2 = fused_param_rows(like.waveform_gen, q, 0.0, 1.0, like.device)
        eager = like.get_ll_fused(q, LENGTHS, modes=MODES, plan=plan, precision=1)  # eager work between replays
        torch.cuda.synchronize()
        s_rows.copy_(r2)
        s_ts.copy_(t2)
        s_te.copy_(e2)
        g.replay()
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        print("fp32 replay", rep, got, eager, flush=True)
        assert int(flags.sum().item()) == 0
        assert np.max(np.abs(got - eager)) < 1e-9
