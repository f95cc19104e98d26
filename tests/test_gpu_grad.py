"""Gradient of the heterodyned log-likelihood (adjoint kernel) against central
finite differences of the FP64 CPU oracle's heterodyned log-likelihood.

The oracle is evaluated at x +- h e_k and x +- 2h e_k (fourth-order central
difference) in the sampling coordinates (ln M, q, chi1z, chi2z, distance,
phi_ref, cos inc, lam, sin beta, psi, t_ref). The oracle's logL carries a
round-off of about 3e-13 |logL| (sums of terms up to 1e6 at |logL| ~ 1e5), so
the steps are chosen large enough that this noise, amplified by 1/h, stays
well below the tolerance, and small enough that the O(h^4) truncation does too
(checked by varying h over four decades): |g_gpu - g_fd| <= 1e-5 |g_fd| per
component.
"""

import numpy as np
import pytest

from oracle.pipeline import OracleHeterodyne, OracleWaveform
from tests.params import MODES, YRSID_SI, draw

pytestmark = pytest.mark.gpu

# finite-difference steps per sampling coordinate
STEPS = np.array([1e-6, 1e-5, 1e-5, 1e-5, 0.0, 1e-5, 1e-5, 1e-5, 1e-5, 1e-5, 1e-1])
REL_DIST_STEP = 1e-5


def to_sampling(p):
    """(12, B) parameter rows -> (B, 11) sampling coordinates."""
    m1, m2 = p[0], p[1]
    return np.stack([np.log(m1 + m2), m1 / m2, p[2], p[3], p[4], p[5], np.cos(p[7]), p[8], np.sin(p[9]), p[10],
                     p[11]], axis=1)


# TDI forms: the benchmark's TDI-2 A/E/T, and X/Y/Z and rescaled first-generation
# A/E/T, whose TDI combinations (and so the adjoint weights of the link terms)
# differ. The latter two put the reference at 100 times the distance: weighted
# by the A/E/T noise model their channels are so loud (|logL| ~ 1e7) that FP64
# rounding alone would exceed the logL tolerances below.
TDI_FORMS = [("AET", True, False), ("XYZ", False, False), ("AET", False, True)]


@pytest.fixture(scope="module", params=TDI_FORMS, ids=["AET-tdi2", "XYZ-tdi1", "AET-tdi1-rescaled"])
def grad_problem(request):
    TDItag, tdi2, rescaled = request.param
    from bbhx_amd import get_backend
    from bbhx_amd.likelihood import HeterodynedLikelihood
    from bbhx_amd.utils.orbits import EqualArmlengthOrbits
    from bbhx_amd.utils.sensitivity import AET1SensitivityMatrix
    from bbhx_amd.waveformbuild import BBHWaveformFD

    backend = get_backend("hip")
    orbits = EqualArmlengthOrbits(dt=1e4, t_max=3.0 * YRSID_SI)
    ref = draw(1, 1001)[:, 0]
    ref[0], ref[1] = 1.3e6, 4.1e5
    ref[11] = 0.62 * YRSID_SI
    if (TDItag, tdi2, rescaled) != TDI_FORMS[0]:
        ref[4] *= 100.0
    f = np.arange(1, 2**15 + 1) * (4.0 / YRSID_SI)
    owave = OracleWaveform(orbits.as_host_dict(), TDItag=TDItag, tdi2=tdi2, rescaled=rescaled)
    inj = owave(*ref, freqs=f, length=1024, modes=MODES, direct=False, fill=True, squeeze=True)[0]
    sens = AET1SensitivityMatrix(f)
    psd = np.array([sens[0], sens[1], sens[2]])
    rng = np.random.default_rng(1001)
    df = f[1] - f[0]
    data = inj + np.sqrt(psd / (4 * df)) * (rng.standard_normal(psd.shape) + 1j * rng.standard_normal(psd.shape))
    K = 96
    wave = BBHWaveformFD(response_kwargs=dict(orbits=orbits, TDItag=TDItag, tdi2=tdi2, rescaled=rescaled),
                         force_backend=backend)
    het = HeterodynedLikelihood(wave, f, data, ref, K, sens_mat=sens, template_gen_kwargs=dict(modes=MODES),
                                reference_gen_kwargs=dict(length=1024, modes=MODES), force_backend=backend)
    psd_fn = lambda x: [AET1SensitivityMatrix(x)[i] for i in range(3)]  # noqa: E731
    ohet = OracleHeterodyne(owave, f, data, ref, K, psd_fn, reference_length=1024, modes=MODES)
    # templates near the reference, far enough that every gradient component is sizable
    x0 = to_sampling(ref[:, None])[0]
    rng = np.random.default_rng(5)
    dx = np.array([3e-5, 1e-3, 3e-3, 3e-3, 0.0, 0.05, 0.02, 0.01, 0.01, 0.05, 3.0])
    X = x0[None, :] + dx[None, :] * rng.standard_normal((3, 11))
    X[:, 4] = x0[4] * (1 + 0.01 * rng.standard_normal(3))
    return dict(het=het, ohet=ohet, X=X)


def oracle_ll(ohet, X):
    from bbhx_amd.likelihood import HeterodynedLikelihood

    return ohet.get_ll(HeterodynedLikelihood.sampling_to_params(X))[0]


def test_grad_value_matches_forward(grad_problem):
    het, ohet, X = grad_problem["het"], grad_problem["ohet"], grad_problem["X"]
    ll_g, grad = het.get_ll_and_grad(X)
    ll_f = het.get_ll(het.sampling_to_params(X), fused=True)
    ll_o = oracle_ll(ohet, X)
    print("logL gradient kernel", ll_g, "forward", ll_f, "oracle", ll_o, flush=True)
    assert np.all(np.isfinite(grad))
    assert np.max(np.abs(ll_g - ll_f)) < 1e-7
    assert np.max(np.abs(ll_g - ll_o)) < 1e-6
    # the logL the gradient entry point writes is the same expression of its d_h, h_h
    het.get_ll_and_grad(X)
    dh, hh = het.hdyn_d_h.cpu().numpy(), het.hdyn_h_h.cpu().numpy()
    assert np.array_equal(ll_g, -1 / 2.0 * (het.reference_d_d + hh - 2 * dh).real)


def test_grad_against_oracle_finite_differences(grad_problem):
    het, ohet, X = grad_problem["het"], grad_problem["ohet"], grad_problem["X"]
    B = X.shape[0]
    _, grad = het.get_ll_and_grad(X)
    steps = np.tile(STEPS, (B, 1))
    steps[:, 4] = REL_DIST_STEP * X[:, 4]
    pts = []
    for k in range(11):
        for s in (2, 1, -1, -2):
            Y = X.copy()
            Y[:, k] += s * steps[:, k]
            pts.append(Y)
    ll = oracle_ll(ohet, np.concatenate(pts)).reshape(11, 4, B)
    llx = oracle_ll(ohet, X)
    fd = (-ll[:, 0] + 8 * ll[:, 1] - 8 * ll[:, 2] + ll[:, 3]) / (12 * steps.T)  # (11, B)
    fd = fd.T
    # bound on the round-off of the difference: the oracle's logL noise (at most 3e-13 |logL|) times ~1.5 / h
    roundoff = 1.5 * 3e-13 * np.maximum(np.abs(llx), 1.0)[:, None] / steps
    err = np.abs(grad - fd)
    rel = err / np.maximum(np.abs(fd), 1e-300)
    names = ("lnM", "q", "chi1", "chi2", "dist", "phi", "cosi", "lam", "sinb", "psi", "tref")
    for b in range(B):
        print(f"template {b}: logL = {llx[b]:.6e}")
        for k in range(11):
            print(f"  {names[k]:5s} gpu {grad[b, k]: .9e} fd {fd[b, k]: .9e} rel {rel[b, k]:.2e} "
                  f"roundoff {roundoff[b, k]:.1e}", flush=True)
    ok = err <= 1e-5 * np.abs(fd)
    assert np.all(ok), np.argwhere(~ok)


def test_autograd_function_backward(grad_problem):
    import torch

    het, X = grad_problem["het"], grad_problem["X"]
    from bbhx_amd.likelihood import sampling_rows_device

    # on the device: the same rows as the device path, so the same gradient bit for bit
    x = torch.tensor(X, dtype=torch.float64, device="cuda", requires_grad=True)
    ll = het.loglike(x)
    w = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64, device="cuda")
    (ll * w).sum().backward()
    ll_d, grad_d = het.get_ll_grad_device(*sampling_rows_device(x.detach()))
    np.testing.assert_array_equal(ll.detach().cpu().numpy(), ll_d.cpu().numpy())
    np.testing.assert_allclose(x.grad.cpu().numpy(), w.cpu().numpy()[:, None] * grad_d.T.cpu().numpy(), rtol=1e-15,
                               atol=0)
    # a CPU tensor gives the same; the host path (numpy parameter rows) agrees to rounding
    xc = torch.tensor(X, dtype=torch.float64, requires_grad=True)
    (het.loglike(xc) * w.cpu()).sum().backward()
    np.testing.assert_allclose(xc.grad.numpy(), x.grad.cpu().numpy(), rtol=1e-15, atol=0)
    _, grad = het.get_ll_and_grad(X)
    np.testing.assert_allclose(x.grad.cpu().numpy(), w.cpu().numpy()[:, None] * grad, rtol=1e-8, atol=0)


def test_grad_graph_replays(grad_problem):
    """The gradient entry point (two streams joined by events inside the
    library) captured in a hipGraph: replays with new inputs match eager calls."""
    import torch

    het, X = grad_problem["het"], grad_problem["X"]
    kw = het._merge_kwargs({})
    rows, ts, te = het.fused_inputs(het.sampling_to_params(X), kw)
    s_rows, s_ts, s_te = rows.clone(), ts.clone(), te.clone()
    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(st):
        het.get_ll_grad_device(s_rows, s_ts, s_te)
    torch.cuda.current_stream().wait_stream(st)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=st):  # the warm-up stream: the library's scratch exists there
        ll_s, grad_s = het.get_ll_grad_device(s_rows, s_ts, s_te)
    torch.cuda.synchronize()
    rng = np.random.default_rng(9)
    for rep in range(3):
        Y = X.copy()
        Y[:, 5] += 0.05 * rng.standard_normal(Y.shape[0])
        Y[:, 10] += 2.0 * rng.standard_normal(Y.shape[0])
        r2, t2, e2 = het.fused_inputs(het.sampling_to_params(Y), kw)
        ll_e, grad_e = het.get_ll_grad_device(r2, t2, e2)  # eager work between replays
        ll_e, grad_e = ll_e.cpu().numpy(), grad_e.cpu().numpy()
        s_rows.copy_(r2)
        s_ts.copy_(t2)
        s_te.copy_(e2)
        g.replay()
        torch.cuda.synchronize()
        print("replay", rep, ll_s.cpu().numpy(), ll_e, flush=True)
        assert np.array_equal(ll_s.cpu().numpy(), ll_e)
        assert np.array_equal(grad_s.cpu().numpy(), grad_e)


def test_grad_graph_from_sampling_coordinates(grad_problem):
    """The device-side inputs from sampling coordinates (sampling_rows_device:
    torch operations and the window kernel) and the gradient, captured in one
    hipGraph: replays with new coordinates copied into the captured input match
    eager calls bit for bit, with eager work between the replays."""
    import torch

    from bbhx_amd.likelihood import sampling_rows_device

    het, X = grad_problem["het"], grad_problem["X"]
    xd = torch.as_tensor(np.ascontiguousarray(X), dtype=torch.float64, device="cuda")
    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(st):
        het.get_ll_grad_device(*sampling_rows_device(xd))
    torch.cuda.current_stream().wait_stream(st)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=st):
        ll_s, grad_s = het.get_ll_grad_device(*sampling_rows_device(xd))
    torch.cuda.synchronize()
    rng = np.random.default_rng(19)
    for rep in range(3):
        Y = X.copy()
        Y[:, 5] += 0.05 * rng.standard_normal(Y.shape[0])
        Y[:, 7] += 0.01 * rng.standard_normal(Y.shape[0])
        Y[:, 10] += 2.0 * rng.standard_normal(Y.shape[0])
        yd = torch.as_tensor(Y, dtype=torch.float64, device="cuda")
        ll_e, grad_e = het.get_ll_grad_device(*sampling_rows_device(yd))  # eager work between replays
        ll_e, grad_e = ll_e.cpu().numpy(), grad_e.cpu().numpy()
        xd.copy_(yd)
        g.replay()
        torch.cuda.synchronize()
        print("replay", rep, ll_s.cpu().numpy()[:3], ll_e[:3], flush=True)
        assert np.array_equal(ll_s.cpu().numpy(), ll_e)
        assert np.array_equal(grad_s.cpu().numpy(), grad_e)


@pytest.mark.parametrize("modes", [[(2, 2)], [(4, 3), (2, 1), (3, 3), (4, 4)]], ids=["22", "43-21-33-44"])
def test_grad_mode_lists(modes):
    """The gradient with other mode lists (one mode; four modes listed out of
    order, one pair of equal m apart) against oracle finite differences, and its
    logL against the forward's."""
    from bbhx_amd import get_backend
    from bbhx_amd.likelihood import HeterodynedLikelihood
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
    inj = owave(*ref, freqs=f, length=1024, modes=modes, direct=False, fill=True, squeeze=True)[0]
    sens = AET1SensitivityMatrix(f)
    psd = np.array([sens[0], sens[1], sens[2]])
    rng = np.random.default_rng(1002)
    df = f[1] - f[0]
    data = inj + np.sqrt(psd / (4 * df)) * (rng.standard_normal(psd.shape) + 1j * rng.standard_normal(psd.shape))
    K = 96
    wave = BBHWaveformFD(response_kwargs=dict(orbits=orbits, tdi2=True), force_backend=backend)
    het = HeterodynedLikelihood(wave, f, data, ref, K, sens_mat=sens, template_gen_kwargs=dict(modes=modes),
                                reference_gen_kwargs=dict(length=1024, modes=modes), force_backend=backend)
    psd_fn = lambda x: [AET1SensitivityMatrix(x)[i] for i in range(3)]  # noqa: E731
    ohet = OracleHeterodyne(owave, f, data, ref, K, psd_fn, reference_length=1024, modes=modes)
    x0 = to_sampling(ref[:, None])[0]
    rng = np.random.default_rng(6)
    dx = np.array([3e-5, 1e-3, 3e-3, 3e-3, 0.0, 0.05, 0.02, 0.01, 0.01, 0.05, 3.0])
    X = x0[None, :] + dx[None, :] * rng.standard_normal((2, 11))
    X[:, 4] = x0[4] * (1 + 0.01 * rng.standard_normal(2))
    ll_g, grad = het.get_ll_and_grad(X)
    ll_f = het.get_ll(het.sampling_to_params(X), fused=True)
    assert np.max(np.abs(ll_g - ll_f)) < 1e-7
    assert np.max(np.abs(ll_g - oracle_ll(ohet, X))) < 1e-6
    B = X.shape[0]
    steps = np.tile(STEPS, (B, 1))
    steps[:, 4] = REL_DIST_STEP * X[:, 4]
    pts = []
    for k in range(11):
        for s in (2, 1, -1, -2):
            Y = X.copy()
            Y[:, k] += s * steps[:, k]
            pts.append(Y)
    ll = oracle_ll(ohet, np.concatenate(pts)).reshape(11, 4, B)
    fd = ((-ll[:, 0] + 8 * ll[:, 1] - 8 * ll[:, 2] + ll[:, 3]) / (12 * steps.T)).T
    rel = np.abs(grad - fd) / np.maximum(np.abs(fd), 1e-300)
    print(modes, "max relative gradient error", rel.max(), flush=True)
    assert np.all(np.abs(grad - fd) <= 1e-5 * np.abs(fd)), np.argwhere(np.abs(grad - fd) > 1e-5 * np.abs(fd))
