"""CPU tests of the host-side raise points (no GPU calls).

They mirror the reference's checks: the response buffer length and frequency
shapes (fastfdresponse.py:389-421), the orbit coverage of t(f)
(fastfdresponse.py:486-492, a debugger stop in the reference, a ValueError
here), the spline shape rules (utils/interpolate.py:93-126), the observation
window search (waveforms/phenomhm.py:520-600) and the size checks of the fused
entry points' bindings (which run before any device call).
"""

from types import SimpleNamespace

import numpy as np
import pytest
import torch

from bbhx_amd.response.fastfdresponse import LISATDIResponse
from bbhx_amd.utils.interpolate import CubicSplineInterpolant
from bbhx_amd.waveforms.phenomhm import PhenomHMAmpPhase


def test_response_buffer_layout():
    B, M, L = 3, 6, 10
    assert LISATDIResponse.buffer_layout(8 * B * M * L, B, M, L) == (8, 0)
    assert LISATDIResponse.buffer_layout(9 * B * M * L, B, M, L) == (9, 1)
    for bad in (7 * B * M * L, 10 * B * M * L, 8 * B * M * L + 1, 0):
        with pytest.raises(ValueError, match="out_buffer incorrect length"):
            LISATDIResponse.buffer_layout(bad, B, M, L)


def test_response_frequency_shapes():
    LISATDIResponse.check_freqs((3, 6, 10), 180, 3, 6, 10)
    LISATDIResponse.check_freqs((180,), 180, 3, 6, 10)
    with pytest.raises(ValueError, match="incorrect shape"):
        LISATDIResponse.check_freqs((3, 5, 12), 180, 3, 6, 10)
    with pytest.raises(ValueError, match="incorrect length"):
        LISATDIResponse.check_freqs((179,), 179, 3, 6, 10)


def test_response_orbit_coverage():
    LISATDIResponse.check_orbit_coverage(0.0, 1e7, 0.0, 1e8)
    with pytest.raises(ValueError, match="does not cover"):
        LISATDIResponse.check_orbit_coverage(-1.0, 1e7, 0.0, 1e8)
    with pytest.raises(ValueError, match="does not cover"):
        LISATDIResponse.check_orbit_coverage(0.0, 2e8, 0.0, 1e8)


def test_spline_shape_rules():
    dims = CubicSplineInterpolant.spline_dims
    assert dims(3, 2 * 6 * 5, (2, 6, 5), 4, 9 * 2 * 6 * 5, (9, 2, 6, 5)) == (9, 2, 6, 5)
    assert dims(1, 60, (60,), 1, 540, (540,), 9, 2, 6, 5) == (9, 2, 6, 5)
    with pytest.raises(ValueError, match="both be flattened"):
        dims(1, 60, (60,), 4, 540, (9, 2, 6, 5))
    with pytest.raises(ValueError, match="dimensional information"):
        dims(1, 60, (60,), 1, 540, (540,), 9, None, 6, 5)
    with pytest.raises(ValueError, match="x array is not correct"):
        dims(1, 61, (61,), 1, 540, (540,), 9, 2, 6, 5)
    with pytest.raises(ValueError, match="y_all array is not correct"):
        dims(1, 60, (60,), 1, 541, (541,), 9, 2, 6, 5)
    with pytest.raises(ValueError, match="3D"):
        dims(2, 60, (12, 5), 4, 540, (9, 2, 6, 5))
    with pytest.raises(ValueError, match="4D"):
        dims(3, 60, (2, 6, 5), 3, 540, (18, 6, 5))
    with pytest.raises(ValueError, match="does not extend"):
        dims(3, 60, (2, 6, 5), 4, 540, (9, 2, 5, 6))
    with pytest.raises(ValueError, match="at least 3 knots"):
        dims(3, 24, (2, 6, 2), 4, 216, (9, 2, 6, 2))


def test_window_bounds_linear_crossings():
    """t(f) linear in f on the grid: the refit ends are the exact crossings of
    t = 0 and t = Tobs; grids inside the window keep their own ends."""
    f = torch.linspace(1e-4, 1e-2, 50, dtype=torch.float64)
    tf = torch.stack([
        (f - 2e-3) * 1e9,          # crosses 0 at f = 2e-3 and Tobs = 5e6 at f = 7e-3
        (f + 1e-3) * 1e8,          # starts at t > 0 and ends below Tobs: no cut
    ])[:, None, :].expand(2, 3, 50).contiguous()
    fs = f.expand(2, 3, 50).contiguous()
    f_min, f_max = PhenomHMAmpPhase.window_bounds(tf, fs, torch.full((2, 3), 5e6, dtype=torch.float64))
    np.testing.assert_allclose(f_min[0].numpy(), 2e-3, rtol=1e-12)
    np.testing.assert_allclose(f_max[0].numpy(), 7e-3, rtol=1e-12)
    np.testing.assert_array_equal(f_min[1].numpy(), 1e-4)
    np.testing.assert_array_equal(f_max[1].numpy(), 1e-2)


def test_window_bounds_double_crossing_raises():
    f = torch.linspace(1e-4, 1e-2, 50, dtype=torch.float64)
    tf = torch.cos(f / 1e-2 * 4 * np.pi)[None, None, :].clone()  # crosses 0 several times
    with pytest.raises(ValueError, match="x_min did not run correctly"):
        PhenomHMAmpPhase.window_bounds(tf, f[None, None, :], torch.full((1, 1), 1e300, dtype=torch.float64))
    with pytest.raises(ValueError, match="x_max did not run correctly"):
        PhenomHMAmpPhase.window_bounds(tf + 2.0, f[None, None, :], torch.full((1, 1), 2.0, dtype=torch.float64))


def test_tobs_grid_forms():
    g = PhenomHMAmpPhase.tobs_grid
    assert torch.all(g(None, 2, 3, "cpu") == 1e300)
    assert torch.all(g(5.0, 2, 3, "cpu") == 5.0)
    np.testing.assert_array_equal(g(np.array([1.0, 2.0]), 2, 3, "cpu").numpy(), [[1.0] * 3, [2.0] * 3])
    np.testing.assert_array_equal(g(np.ones((2, 3)), 2, 3, "cpu").numpy(), np.ones((2, 3)))
    for bad in (np.ones(3), np.ones((3, 2)), 5, "x"):
        with pytest.raises(ValueError, match="Tobs entered incorrectly"):
            g(bad, 2, 3, "cpu")


def _hdyn_args(B=4, K=8, **over):
    z = lambda *s, dt=torch.float64: torch.zeros(*s, dtype=dt)  # noqa: E731
    a = dict(d_h=z(B, dt=torch.complex128), h_h=z(B, dt=torch.complex128), params=z(11, B), t_start=z(B),
             t_end=z(B), ells=np.array([2], np.int32), mms=np.array([2], np.int32), freqs=z(K),
             h0=z(3, K, dt=torch.complex128), data_constants=z(12 * K, dt=torch.complex128), qnm=z(8 * 1001), dspin=0.002,
             initial_t_val=0.0, response=None, TDItag=2, rescaled=False, tdi2=True)
    a.update(over)
    return a


@pytest.mark.parametrize("field,bad", [
    ("h0", torch.zeros(3, 7, dtype=torch.complex128)),
    ("data_constants", torch.zeros(12 * 8 - 1, dtype=torch.complex128)),
    ("d_h", torch.zeros(3, dtype=torch.complex128)),
    ("h_h", torch.zeros(5, dtype=torch.complex128)),
    ("t_start", torch.zeros(3, dtype=torch.float64)),
    ("t_end", torch.zeros(5, dtype=torch.float64)),
    ("params", torch.zeros(12, 4, dtype=torch.float64)),
    ("qnm", torch.zeros(15, dtype=torch.float64)),
])
def test_fused_bindings_check_sizes_on_the_host(field, bad):
    from bbhx_amd import cutils

    with pytest.raises(ValueError):
        cutils.hdyn_fused(**_hdyn_args(**{field: bad}))
    with pytest.raises(ValueError):
        cutils.hdyn_fused_grad(grad=torch.zeros(11, 4, dtype=torch.float64), **_hdyn_args(**{field: bad}))
    with pytest.raises(ValueError):
        cutils.hdyn_fused(**_hdyn_args(), logL=torch.zeros(3, dtype=torch.float64))


@pytest.mark.parametrize("field,bad", [("flags", torch.zeros(3, dtype=torch.int32)),
                                       ("data_freqs", torch.zeros(99, dtype=torch.float64)),
                                       ("data_channels", torch.zeros(3, 99, dtype=torch.complex128)),
                                       ("noise_factors", torch.zeros(3, 101, dtype=torch.float64)),
                                       ("d_h", torch.zeros(5, dtype=torch.complex128))])
def test_direct_fused_binding_checks_sizes_on_the_host(field, bad):
    from bbhx_amd import cutils

    B, N = 4, 100
    plan = SimpleNamespace(num_bin_all=B, data_length=N, handle=None, chunk=4096, block=256, grid=64)
    z = lambda *s, dt=torch.float64: torch.zeros(*s, dtype=dt)  # noqa: E731
    a = dict(d_h=z(B, dt=torch.complex128), h_h=z(B, dt=torch.complex128), flags=z(B, dt=torch.int32),
             params=z(11, B), t_start=z(B), t_end=z(B), ells=np.array([2], np.int32), mms=np.array([2], np.int32),
             data_freqs=z(N), data_channels=z(3, N, dt=torch.complex128), noise_factors=z(3, N), qnm=z(8 * 1001),
             dspin=0.002, initial_t_val=0.0, mf_min=1e-4, mf_max=0.6, response=None, TDItag=2, rescaled=False,
             tdi2=True)
    a[field] = bad
    with pytest.raises(ValueError):
        cutils.direct_fused(plan, **a)
    with pytest.raises(ValueError):
        cutils.direct_fused_grad(plan, grad=z(11, B), **a)


def test_oracle_under_address_and_ub_sanitizers():
    """`make -C oracle sanitize`: every oracle entry point on a small problem,
    built with AddressSanitizer and UBSan (host code only)."""
    import shutil
    import subprocess
    from pathlib import Path

    if shutil.which("g++") is None:
        pytest.skip("no host compiler")
    root = Path(__file__).resolve().parent.parent
    r = subprocess.run(["make", "-C", str(root / "oracle"), "sanitize"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "sanitize driver: OK" in r.stdout
