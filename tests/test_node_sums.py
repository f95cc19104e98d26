"""Per-node form of the heterodyned sums (CPU only).

The fused heterodyned kernels sum per node instead of per segment: with
W_n, V_n, U_n formed once from the heterodyne constants (csrc/fused_device.h,
NodeW and node_weight), d_h = sum conj(r_n) W_n and h_h = sum |r_n|^2 V_n +
Re(r_n conj(r_{n-1})) U_n. This checks that identity, restated here in numpy,
against the oracle's segment sums (Likelihood.cu:52-111 restated in
oracle/oracle_core.cpp) on random residuals and constants, including a first
node at f = 0 (its segment is skipped) and both last-segment rules, so that a
wrong weight, a dropped neighbour term or a sign in the cross term fails.
"""

import numpy as np
import pytest

import oracle


def node_form(r, dc, freqs, jend):
    """(d_h, h_h) of r (B, 3, K) complex from dc (4, 3, K) by the per-node weights."""
    K = len(freqs)
    A0, A1, B0, B1 = dc
    W = np.zeros((3, K), complex)
    V = np.zeros((3, K), complex)
    U = np.zeros((3, K), complex)
    valid = [1 <= j < jend and freqs[j - 1] != 0.0 for j in range(K)]
    for n in range(K):
        if valid[n]:
            inv = 1.0 / (freqs[n] - freqs[n - 1])
            W[:, n] += 0.5 * A0[:, n] + A1[:, n] * inv
            V[:, n] += 0.25 * B0[:, n] + B1[:, n] * inv
            U[:, n] = 0.5 * B0[:, n]
        if n + 1 < K and valid[n + 1]:
            inv = 1.0 / (freqs[n + 1] - freqs[n])
            W[:, n] += 0.5 * A0[:, n + 1] - A1[:, n + 1] * inv
            V[:, n] += 0.25 * B0[:, n + 1] - B1[:, n + 1] * inv
    d_h = np.sum(np.conj(r) * W, axis=(1, 2))
    cross = np.zeros(r.shape)
    cross[:, :, 1:] = (r[:, :, 1:] * np.conj(r[:, :, :-1])).real
    h_h = np.sum(np.abs(r) ** 2 * V + cross * U, axis=(1, 2))
    return d_h, h_h


@pytest.mark.parametrize("zero_first", [False, True])
@pytest.mark.parametrize("include_last", [False, True])
def test_node_form_equals_segment_sums(zero_first, include_last):
    rng = np.random.default_rng(7 + 2 * zero_first + include_last)
    B, K = 5, 40
    freqs = np.sort(rng.uniform(1e-4, 1e-1, K))
    if zero_first:
        freqs[0] = 0.0
    r = rng.standard_normal((B, 3, K)) + 1j * rng.standard_normal((B, 3, K))
    dc = rng.standard_normal((4, 3, K)) + 1j * rng.standard_normal((4, 3, K))
    res = r.transpose((2, 1, 0)).flatten()
    d_seg, h_seg = oracle.hdyn(res, dc.flatten(), freqs, B, K, include_last_segment=include_last)
    d_node, h_node = node_form(r, dc, freqs, K if include_last else K - 1)
    scale_d = np.max(np.abs(d_seg))
    scale_h = np.max(np.abs(h_seg))
    assert np.max(np.abs(d_node - d_seg)) <= 1e-12 * scale_d
    assert np.max(np.abs(h_node - h_seg)) <= 1e-12 * scale_h


def test_node_form_reference_identity():
    """r = 1 at every node gives d_h = sum A0 and h_h = sum B0 (Likelihood.cu:91-103
    with r0 = 1, r1 = 0)."""
    rng = np.random.default_rng(3)
    K = 32
    freqs = np.sort(rng.uniform(1e-4, 1e-1, K))
    dc = rng.standard_normal((4, 3, K)) + 1j * rng.standard_normal((4, 3, K))
    r = np.ones((1, 3, K), complex)
    d_h, h_h = node_form(r, dc, freqs, K - 1)
    assert abs(d_h[0] - dc[0][:, 1:K - 1].sum()) <= 1e-12 * np.abs(dc[0]).sum()
    assert abs(h_h[0] - dc[2][:, 1:K - 1].sum()) <= 1e-12 * np.abs(dc[2]).sum()


def node_adjoint(r, dc, freqs, jend):
    """gamma (B, 3, K) with d logL = sum Re(conj(gamma) dr), logL = Re d_h - Re h_h / 2,
    from the per-node weights (csrc/hdyn_grad.hip, adjoint_weights)."""
    K = len(freqs)
    A0, A1, B0, B1 = dc
    W = np.zeros((3, K), complex)
    V = np.zeros((3, K))
    U = np.zeros((3, K))
    valid = [1 <= j < jend and freqs[j - 1] != 0.0 for j in range(K)]
    for n in range(K):
        if valid[n]:
            inv = 1.0 / (freqs[n] - freqs[n - 1])
            W[:, n] += 0.5 * A0[:, n] + A1[:, n] * inv
            V[:, n] += (0.25 * B0[:, n] + B1[:, n] * inv).real
            U[:, n] = (0.5 * B0[:, n]).real
        if n + 1 < K and valid[n + 1]:
            inv = 1.0 / (freqs[n + 1] - freqs[n])
            W[:, n] += 0.5 * A0[:, n + 1] - A1[:, n + 1] * inv
            V[:, n] += (0.25 * B0[:, n + 1] - B1[:, n + 1] * inv).real
    gam = W[None] - V[None] * r
    gam[:, :, 1:] -= 0.5 * U[None, :, 1:] * r[:, :, :-1]
    gam[:, :, :-1] -= 0.5 * U[None, :, 1:] * r[:, :, 1:]
    return gam


@pytest.mark.parametrize("zero_first", [False, True])
@pytest.mark.parametrize("include_last", [False, True])
def test_node_adjoint_matches_segment_differences(zero_first, include_last):
    """The adjoint weights of the per-node form against central differences of
    the oracle's segment sums (logL is quadratic in r, so the differences are
    exact up to rounding), along random directions and along single nodes at
    both ends, where a missing neighbour term or a wrong end rule shows."""
    rng = np.random.default_rng(11 + 2 * zero_first + include_last)
    B, K = 3, 24
    freqs = np.sort(rng.uniform(1e-4, 1e-1, K))
    if zero_first:
        freqs[0] = 0.0
    r = rng.standard_normal((B, 3, K)) + 1j * rng.standard_normal((B, 3, K))
    dc = rng.standard_normal((4, 3, K)) + 1j * rng.standard_normal((4, 3, K))

    def loglike(rr):
        res = rr.transpose((2, 1, 0)).flatten()
        d_h, h_h = oracle.hdyn(res, dc.flatten(), freqs, B, K, include_last_segment=include_last)
        return d_h.real - 0.5 * h_h.real

    gam = node_adjoint(r, dc, freqs, K if include_last else K - 1)
    dirs = [rng.standard_normal((B, 3, K)) + 1j * rng.standard_normal((B, 3, K)) for _ in range(3)]
    for n in (0, 1, K - 2, K - 1):
        for c in range(3):
            for unit in (1.0, 1j):
                e = np.zeros((B, 3, K), complex)
                e[:, c, n] = unit
                dirs.append(e)
    eps = 1e-3
    for dr in dirs:
        fd = (loglike(r + eps * dr) - loglike(r - eps * dr)) / (2 * eps)
        an = np.sum((np.conj(gam) * dr).real, axis=(1, 2))
        scale = np.sum(np.abs(gam) * np.abs(dr), axis=(1, 2)) + 1e-300
        assert np.max(np.abs(fd - an) / scale) <= 1e-9, (dr.nonzero(), fd, an)
