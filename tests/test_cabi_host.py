"""CPU tests of the C ABI library and of the host-side logic (no GPU calls)."""

import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "bbhx_hip.h"
LIB = ROOT / "bbhx_amd" / "libbbhx_hip.so"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(bbhx_hip_\w+)\s*\(", text)))


def test_header_declares_the_backend_methods():
    names = declared_symbols()
    for n in ("bbhx_hip_waveform_amp_phase", "bbhx_hip_get_phenomhm_ringdown_frequencies",
              "bbhx_hip_get_phenomd_ringdown_frequencies", "bbhx_hip_response_create",
              "bbhx_hip_response_add_orbits", "bbhx_hip_response_compute", "bbhx_hip_response_destroy",
              "bbhx_hip_interpolate", "bbhx_hip_InterpTDI", "bbhx_hip_direct_sum", "bbhx_hip_hdyn",
              "bbhx_hip_direct_like", "bbhx_hip_hdyn_fused", "bbhx_hip_hdyn_fused_grad", "bbhx_hip_direct_plan_create",
              "bbhx_hip_direct_plan_destroy", "bbhx_hip_direct_plan_band", "bbhx_hip_direct_fused",
              "bbhx_hip_hdyn_prep", "bbhx_hip_direct_plan_enable_grad", "bbhx_hip_direct_fused_grad",
              "bbhx_hip_direct_plan_set_precision"):
        assert n in names, n


@pytest.mark.skipif(not LIB.exists(), reason="libbbhx_hip.so not built")
def test_library_loads_and_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    for name in declared_symbols():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (bbhx_hip_\w+)", out))
    assert set(declared_symbols()) <= exported
    from bbhx_amd.cutils import exported_symbols

    assert set(exported_symbols()) == set(declared_symbols())


@pytest.mark.skipif(not LIB.exists(), reason="libbbhx_hip.so not built")
def test_version_and_argument_errors_need_no_gpu():
    """Entry points validate sizes before touching the device and return status 1."""
    from bbhx_amd.cutils import load_library

    lib = load_library()
    assert b"gfx950" in lib.bbhx_hip_version()
    # negative sizes / null pointers are rejected on the host
    assert lib.bbhx_hip_interpolate(None, None, None, None, None, None, 2, 9, 6, 1, None) == 1
    assert lib.bbhx_hip_hdyn(None, None, None, None, None, -1, 10, 3, 1, None) == 1
    assert lib.bbhx_hip_direct_sum(None, None, 1, 10, 2, 6, None, None, None) == 1
    assert lib.bbhx_hip_waveform_amp_phase(None, None, None, None, None, None, None, None, None, None, 7, 10, 1,
                                           None, None, 0, None) == 1
    assert lib.bbhx_hip_response_compute(None, None, None, None, None, None, None, None, None, None, 2, 0, 1, 0, 6,
                                         10, 1, 1, None) == 1
    # ragged direct plan: lengths below 4, bad min_chunk, no lengths
    import ctypes as C

    h = C.c_void_p()
    lens = (C.c_int * 3)(512, 3, 600)
    assert lib.bbhx_hip_direct_plan_create(C.byref(h), 3, lens, 6, 1000, 1024) == 1
    lens_ok = (C.c_int * 3)(512, 300, 600)
    assert lib.bbhx_hip_direct_plan_create(C.byref(h), 3, lens_ok, 6, 1000, 16) == 1
    assert lib.bbhx_hip_direct_plan_create(C.byref(h), 3, None, 6, 1000, 1024) == 1
    assert lib.bbhx_hip_direct_plan_create(C.byref(h), 3, lens_ok, 7, 1000, 1024) == 1
    assert lib.bbhx_hip_direct_fused(None, *([None] * 8), 6, *([None] * 4), 2, 2e-3, 0.0, 1e-4, 0.6, None, 2, 0, 1,
                                     16384, 256, 64, None) == 1
    assert lib.bbhx_hip_direct_plan_bytes(None) == -1
    assert lib.bbhx_hip_direct_plan_band(None, None) == 1
    assert lib.bbhx_hip_direct_plan_enable_grad(None) == 1
    assert lib.bbhx_hip_direct_plan_set_precision(None, 0) == 1
    assert lib.bbhx_hip_direct_plan_set_precision(None, 1) == 1
    # heterodyne constants: too few points or nodes, df <= 0, missing arrays
    lib.bbhx_hip_hdyn_prep.argtypes = [C.c_void_p] * 6 + [C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_void_p]
    assert lib.bbhx_hip_hdyn_prep(None, None, None, None, None, None, 1, None, 8, 1e-8, None) == 1
    assert lib.bbhx_hip_hdyn_prep(None, None, None, None, None, None, 100, None, 1, 1e-8, None) == 1
    assert lib.bbhx_hip_hdyn_prep(None, None, None, None, None, None, 100, None, 8, 0.0, None) == 1
    assert lib.bbhx_hip_hdyn_prep(None, None, None, None, None, None, 100, None, 8, 1e-8, None) == 1
    # window times: missing arrays, negative count
    assert lib.bbhx_hip_window_times(None, None, None, None, 5, 0.0, 1.0, None) == 1
    assert lib.bbhx_hip_window_times(None, None, None, None, -1, 0.0, 1.0, None) == 1


def test_registry_without_gpu_fails_loudly(monkeypatch):
    import torch

    import bbhx_amd

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    bbhx_amd._INSTANCES.clear()
    with pytest.raises(bbhx_amd.BackendUnavailableException):
        bbhx_amd.get_backend("hip")
    assert not bbhx_amd.has_backend("gpu")
    with pytest.raises(bbhx_amd.BackendUnavailableException):
        bbhx_amd.get_backend("cpu")  # there is no CPU fallback


def test_backend_name_aliases():
    import bbhx_amd

    for n in ("hip", "gpu", "cuda", "rocm", "bbhx_hip", "cuda12x", "bbhx_cuda12x"):
        assert bbhx_amd._canonical(n) == "bbhx_hip"


def test_sanity_check_params_swaps_and_rejects():
    from bbhx_amd.waveforms.phenomhm import PhenomHMAmpPhase

    m1 = np.array([1e5, 3e6])
    m2 = np.array([2e5, 1e6])
    c1 = np.array([0.1, 0.2])
    c2 = np.array([-0.3, 0.4])
    m1, m2, c1, c2 = PhenomHMAmpPhase._sanity_check_params(None, m1, m2, c1, c2)
    np.testing.assert_array_equal(m1, [2e5, 3e6])
    np.testing.assert_array_equal(m2, [1e5, 1e6])
    np.testing.assert_array_equal(c1, [-0.3, 0.2])
    np.testing.assert_array_equal(c2, [0.1, 0.4])
    with pytest.raises(ValueError):
        PhenomHMAmpPhase._sanity_check_params(None, np.array([1.0]), np.array([1.0]), np.array([1.2]), np.array([0.]))
    with pytest.raises(ValueError):
        PhenomHMAmpPhase._sanity_check_params(None, np.array([np.nan]), np.array([1.0]), np.array([0.]),
                                              np.array([0.]))


def test_tssb_from_lframe_first_order_and_inverse():
    from bbhx_amd.utils.constants import AU_SI, C_SI
    from bbhx_amd.utils.transform import tLfromSSBframe, tSSBfromLframe

    tL = np.linspace(0, 3e7, 7)
    lam, beta = 0.7, -0.4
    tS = tSSBfromLframe(tL, lam, beta)
    # the delay is bounded by R/c cos(beta)
    assert np.all(np.abs(tS - tL) <= AU_SI / C_SI * np.cos(beta) * (1 + 1e-3))
    # first-order inverse recovers tL to O(Omega (R/c)^2) ~ 5e-2 s
    assert np.max(np.abs(tLfromSSBframe(tS, lam, beta) - tL)) < 0.1


def test_equal_arm_orbit_geometry():
    from bbhx_amd.utils.constants import AU_SI, C_SI, YRSID_SI
    from bbhx_amd.utils.orbits import EqualArmlengthOrbits

    o = EqualArmlengthOrbits(dt=1e5, t_max=1.0 * YRSID_SI)
    x = o.x_base
    centre = x.mean(axis=1)
    r = np.sqrt((centre**2).sum(-1))
    assert np.allclose(r, AU_SI, rtol=1e-3)
    # arms are equal to L to first order in e
    for (i, j) in ((0, 1), (1, 2), (2, 0)):
        d = np.sqrt(((x[:, i] - x[:, j]) ** 2).sum(-1))
        assert np.allclose(d, o.armlength, rtol=3e-2)
    # unit vectors, link ij pointing from j to i, opposite links opposite
    assert np.allclose((o.n_base**2).sum(-1), 1.0)
    li = {l: k for k, l in enumerate(o.LINKS)}
    assert np.allclose(o.n_base[:, li[12]], -o.n_base[:, li[21]])
    assert np.allclose(o.ltt_base * C_SI, np.sqrt(((x[:, [0, 1, 2, 0, 2, 1]] - x[:, [1, 2, 0, 2, 1, 0]]) ** 2).sum(-1)))
    # the constellation plane is tilted by 60 degrees to the ecliptic
    nrm = np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0])
    nrm /= np.sqrt((nrm**2).sum(-1))[:, None]
    assert np.allclose(np.abs(nrm[:, 2]), 0.5, atol=2e-2)


def test_sensitivity_positive_and_t_suppressed():
    from bbhx_amd.utils.sensitivity import AET1SensitivityMatrix, get_sensitivity

    f = np.logspace(-5, -1, 400)
    s = AET1SensitivityMatrix(f)
    assert np.all(s[0] > 0) and np.all(s[2] > 0)
    np.testing.assert_array_equal(s[0], s[1])
    low = f < 1e-3
    assert np.all(s[2][low] < 1e-2 * s[0][low])  # T is insensitive at low frequency
    np.testing.assert_array_equal(get_sensitivity(f, sens_fn="T1TDISens"), s[2])


def test_sampling_coordinates_round_trip():
    from bbhx_amd.likelihood import HeterodynedLikelihood

    rng = np.random.default_rng(4)
    B = 5
    x = np.stack([rng.uniform(11, 16, B), rng.uniform(1, 10, B), rng.uniform(-0.9, 0.9, B), rng.uniform(-0.9, 0.9, B),
                  rng.uniform(1e25, 1e26, B), rng.uniform(0, 6, B), rng.uniform(-0.99, 0.99, B), rng.uniform(0, 6, B),
                  rng.uniform(-0.99, 0.99, B), rng.uniform(0, 3, B), rng.uniform(1e6, 3e7, B)], axis=1)
    p = HeterodynedLikelihood.sampling_to_params(x)
    assert p.shape == (12, B)
    np.testing.assert_allclose(np.log(p[0] + p[1]), x[:, 0], rtol=1e-14)
    np.testing.assert_allclose(p[0] / p[1], x[:, 1], rtol=1e-14)
    np.testing.assert_allclose(np.cos(p[7]), x[:, 6], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(np.sin(p[9]), x[:, 8], rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(p[6], 0.0)
    assert np.all(p[0] >= p[1])


def test_uneven_lengths_rule_spans_range():
    import bench
    from tests.params import draw

    p = draw(8192, 1002, mtot=(1e4, 1e8))
    L = bench.ragged_lengths(p, 31558149.763545603)
    assert L.min() == 256 and L.max() <= 4096 and L.max() > 3000
    # longer grids for wider bands: monotone in the log-span
    assert np.corrcoef(L, np.log(p[0] + p[1]))[0, 1] < 0


def test_qnm_table_file_matches_data_module():
    from bbhx_amd.waveforms import ringdownphenomd as r

    assert r.QNMData_a.size == r.QNMData_fring.size == r.QNMData_fdamp.size == 1001
    np.testing.assert_allclose(np.diff(r.QNMData_a), r.dspin, rtol=1e-9)


def test_sampling_rows_match_host_rows():
    """The device-side parameter rows (torch operations, here on CPU tensors)
    equal the host path's (numpy) rows to rounding, with the m1 >= m2 swap."""
    import torch

    from bbhx_amd.likelihood import HeterodynedLikelihood, sampling_rows

    rng = np.random.default_rng(8)
    B = 6
    x = np.stack([rng.uniform(11, 16, B), rng.uniform(1, 10, B), rng.uniform(-0.9, 0.9, B), rng.uniform(-0.9, 0.9, B),
                  rng.uniform(1e25, 1e26, B), rng.uniform(0, 6, B), rng.uniform(-0.99, 0.99, B), rng.uniform(0, 6, B),
                  rng.uniform(-0.99, 0.99, B), rng.uniform(0, 3, B), rng.uniform(1e6, 3e7, B)], axis=1)
    x[0, 1] = 0.5  # q < 1: the swap
    rows = sampling_rows(torch.as_tensor(x))
    p = HeterodynedLikelihood.sampling_to_params(x)
    m1, m2, c1, c2 = p[0].copy(), p[1].copy(), p[2].copy(), p[3].copy()
    sw = m1 < m2
    m1[sw], m2[sw], c1[sw], c2[sw] = m2[sw], m1[sw], c2[sw], c1[sw]
    want = np.stack([m1, m2, c1, c2, p[4], p[5], p[7], p[8], p[9], p[10], p[11]])
    np.testing.assert_allclose(rows.numpy(), want, rtol=1e-15, atol=0)
    assert rows.numpy()[0, 0] >= rows.numpy()[1, 0]
