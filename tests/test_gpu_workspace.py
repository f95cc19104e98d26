"""The library's scratch buffers (one grow-only set per stream, replaced
buffers kept until release): a graph captured on a stream still replays
correctly after a larger eager call on that stream has grown the buffers it
captured, and calls on two streams do not share buffers (two different
batches evaluated concurrently on two streams give each its own results)."""

import numpy as np
import pytest

from tests.params import MODES, YRSID_SI, draw, perturb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def het_ws():
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
    data = wave(*ref, freqs=f, length=1024, modes=MODES, direct=False, fill=True)[0].cpu().numpy()
    het = HeterodynedLikelihood(wave, f, data, ref, 64, sens_mat=sens, template_gen_kwargs=dict(modes=MODES),
                                reference_gen_kwargs=dict(length=2048, modes=MODES))
    return het, ref


def test_graph_replay_after_workspace_growth(het_ws):
    import torch

    het, ref = het_ws
    p = perturb(ref, 512, 21, scale=1e-3)
    small = het.fused_inputs(p[:, :16])
    big = het.fused_inputs(p)
    eager_small = het.get_ll_device(*small).cpu().numpy()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        het.get_ll_device(*small)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        out = het.get_ll_device(*small)
    g.replay()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), eager_small)
    # a larger batch on the capture stream grows (replaces) the buffers the graph holds
    with torch.cuda.stream(s):
        ll_big = het.get_ll_device(*big)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ll_big.cpu().numpy()[:16], eager_small)
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), eager_small)


def test_two_streams_do_not_share_scratch(het_ws):
    import torch

    het, ref = het_ws
    pa = perturb(ref, 400, 31, scale=1e-3)
    pb = perturb(ref, 300, 32, scale=2e-3)
    ia, ib = het.fused_inputs(pa), het.fused_inputs(pb)
    ea = het.get_ll_device(*ia).cpu().numpy()
    eb = het.get_ll_device(*ib).cpu().numpy()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    sa.wait_stream(torch.cuda.current_stream())
    sb.wait_stream(torch.cuda.current_stream())
    outs = []
    for _ in range(4):
        with torch.cuda.stream(sa):
            outs.append(("a", het.get_ll_device(*ia)))
        with torch.cuda.stream(sb):
            outs.append(("b", het.get_ll_device(*ib)))
    torch.cuda.synchronize()
    for name, o in outs:
        np.testing.assert_array_equal(o.cpu().numpy(), ea if name == "a" else eb)
