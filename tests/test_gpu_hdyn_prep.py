"""Heterodyne constants on the device (bbhx_hip_hdyn_prep) against the FP64
oracle's restatement of likelihood.py:499-533 (oracle.pipeline.hdyn_constants,
pinned by a per-point brute-force loop and the exactness of the heterodyne for a
segment-wise linear ratio in tests/test_oracle_pins.py).

Tolerance: 1e-12 relative to the largest constant of each kind (FP64 on both
sides; the sums run in a different order). The device result is bit-identical
from call to call (fixed summation order)."""

import numpy as np
import pytest

from oracle.pipeline import hdyn_constants

pytestmark = pytest.mark.gpu


def _problem(seed, N, K, empty_low=False):
    rng = np.random.default_rng(seed)
    f = np.arange(1, N + 1) * 3.2e-8
    lo = f[0] * (0.999999 if not empty_low else 0.2)  # nodes denser than the bins at the low end
    nodes = np.logspace(np.log10(lo), np.log10(f[-1] * 1.000001), K)
    keep = (f >= nodes[0]) & (f <= nodes[-1])
    f = f[keep]
    d = rng.standard_normal((3, f.size)) + 1j * rng.standard_normal((3, f.size))
    h0 = rng.standard_normal((3, f.size)) + 1j * rng.standard_normal((3, f.size))
    S = rng.uniform(0.5, 2.0, (3, f.size)) * 1e-40
    return f, nodes, d, h0, S


def _device(f, nodes, d, h0, S):
    import torch

    from bbhx_amd import get_backend

    be = get_backend("hip")
    dev = torch.device("cuda")
    K = nodes.size
    out = torch.zeros(12 * K, dtype=torch.complex128, device=dev)
    tot = torch.zeros(3, dtype=torch.float64, device=dev)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)  # noqa: E731
    be.hdyn_prep(out, tot, t(f), t(d), t(h0), t(S), t(nodes), float(f[1] - f[0]))
    torch.cuda.synchronize()
    return out.cpu().numpy(), tot.cpu().numpy()


@pytest.mark.parametrize("N,K,empty_low", [(5000, 37, False), (70000, 128, False), (3000, 400, True)])
def test_hdyn_prep_matches_oracle(N, K, empty_low):
    f, nodes, d, h0, S = _problem(7 + K, N, K, empty_low)
    print(f"hdyn_prep N={f.size} K={K} empty_low={empty_low}", flush=True)
    got, tot = _device(f, nodes, d, h0, S)
    want, wtot = hdyn_constants(f, d, h0, S, nodes)
    g4, w4 = got.reshape(4, 3, K), want.reshape(4, 3, K)
    for q in range(4):
        scale = np.abs(w4[q]).max()
        assert np.max(np.abs(g4[q] - w4[q])) <= 1e-12 * scale, q
    np.testing.assert_allclose(tot, np.array(wtot), rtol=1e-12)
    if empty_low:
        assert np.sum(np.all(w4[0] == 0, axis=0)) > 1  # some segments hold no dense point


def test_hdyn_prep_deterministic():
    f, nodes, d, h0, S = _problem(3, 40000, 96)
    a, ta = _device(f, nodes, d, h0, S)
    b, tb = _device(f, nodes, d, h0, S)
    assert np.array_equal(a, b) and np.array_equal(ta, tb)
