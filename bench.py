"""Heterodyned log-likelihood throughput on one MI355X.

Workload (the heterodyned-likelihood configuration): one reference MBHB (the
first draw of the seeded parameter distribution), 4096 templates with every
sampled parameter offset by N(0, 1e-3 * prior range), PhenomHM modes
(2,2),(2,1),(3,3),(3,2),(4,4),(4,3), TDI-2 A/E/T response, K = 128
log-spaced sparse nodes, dense data of 2^20 bins at df = 1/yr (injection +
Gaussian noise from the analytic A/E/T PSD), complex128 data, FP64 PSD.

A step = one heterodyned log-likelihood evaluation of all templates with the
parameters already resident in HBM (fused kernel: waveform + response + mode
sum + r = h/h0 + heterodyned sums). Prints ONE JSON line.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--batch 4096] [--nodes 128]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def graph_time(fn, steps, warmup):
    """Capture fn() (device work only, no host synchronisation) in a hipGraph
    after warm-up calls on the capture stream (the library sizes its scratch
    buffers there; it never allocates during a capture); returns (ms per
    replay over `steps` replays, the captured outputs)."""
    import torch

    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(st):
        for _ in range(max(warmup, 1)):
            fn()
    torch.cuda.current_stream().wait_stream(st)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=st):
        out = fn()
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps, out, g


def build_problem(batch, nodes, dense_bits, seed=1000):
    import torch

    from bbhx_amd.likelihood import HeterodynedLikelihood
    from bbhx_amd.utils.constants import YRSID_SI
    from bbhx_amd.utils.orbits import EqualArmlengthOrbits
    from bbhx_amd.utils.sensitivity import AET1SensitivityMatrix
    from bbhx_amd.waveformbuild import BBHWaveformFD
    from tests.params import MODES, draw, perturb

    ref = draw(16, seed)[:, 0]
    templates = perturb(ref, batch, seed + 1, scale=1e-3)
    f = np.arange(1, 2**dense_bits + 1) / YRSID_SI
    orbits = EqualArmlengthOrbits(dt=1e4, t_max=3.0 * YRSID_SI)
    wave = BBHWaveformFD(response_kwargs=dict(orbits=orbits, tdi2=True))
    sens = AET1SensitivityMatrix(f)
    psd = np.array([sens[0], sens[1], sens[2]])
    inj = wave(*ref, freqs=f, length=1024, modes=MODES, direct=False, fill=True)[0].cpu().numpy()
    rng = np.random.default_rng(seed + 1)
    df = f[1] - f[0]
    noise = np.sqrt(psd / (4 * df)) * (rng.standard_normal(psd.shape) + 1j * rng.standard_normal(psd.shape))
    data = inj + noise
    het = HeterodynedLikelihood(wave, f, data, ref, nodes, sens_mat=sens, template_gen_kwargs=dict(modes=MODES),
                                reference_gen_kwargs=dict(length=8096, modes=MODES))
    torch.cuda.synchronize()
    return dict(ref=ref, templates=templates, f=f, data=data, psd=psd, het=het, wave=wave, orbits=orbits)


def _omp_threads(n):
    """Set the OpenMP thread count of the oracle's runtime (libgomp) in this process."""
    import ctypes

    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))
        return True
    except OSError:
        return False


def cpu_baseline(prob, sample, nodes, min_seconds=10.0, one_thread_seconds=4.0):
    """The FP64 CPU oracle (OpenMP over templates on the host cores) on a bounded
    sample: the first `sample` templates of the batch, repeated for at least
    `min_seconds` of CPU time; then the same on one thread (the reference's own
    CPU build is single-threaded) for at least `one_thread_seconds`; then the
    oracle built with FMA contraction on the same sample (the FP64 rounding
    spread of the algorithm itself)."""
    import oracle
    from bbhx_amd.utils.sensitivity import AET1SensitivityMatrix
    from oracle.pipeline import OracleHeterodyne, OracleWaveform
    from tests.params import MODES

    psd_fn = lambda x: [AET1SensitivityMatrix(x)[i] for i in range(3)]  # noqa: E731

    def make():
        owave = OracleWaveform(prob["orbits"].as_host_dict(), tdi2=True)
        return OracleHeterodyne(owave, prob["f"], prob["data"], prob["ref"], nodes, psd_fn, reference_length=8096,
                                modes=MODES)

    ohet = make()
    p = prob["templates"][:, :sample]
    ohet.get_ll(p[:, :8])  # warm-up
    t0 = time.perf_counter()
    reps = 0
    ll = None
    while True:
        ll = ohet.get_ll(p)[0]
        reps += 1
        if time.perf_counter() - t0 > min_seconds:
            break
    dt = (time.perf_counter() - t0) / reps
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    one = None
    if _omp_threads(1):
        n1 = min(256, sample)
        t0 = time.perf_counter()
        r1 = 0
        while True:
            ohet.get_ll(p[:, :n1])
            r1 += 1
            if time.perf_counter() - t0 > one_thread_seconds:
                break
        one = n1 * r1 / (time.perf_counter() - t0)
        _omp_threads(cores)
    with oracle.fma_rounding():
        ll_fma = make().get_ll(p)[0]
    return dict(value=sample / dt, ll=ll, ll_fma=ll_fma, cores=cores, seconds_per_batch=dt, reps=reps,
                value_1thread=one)


def ragged_lengths(p, t_data):
    """Sparse grid length of each template of the uneven batch, proportional to
    the log-span of its observable band: f_lo = the leading-order PN frequency
    a time t_ref before merger (or 1/T_data if higher), f_hi = min(0.1 / M_sec,
    Nyquist 0.05 Hz); L = clip(2400 (log10(f_hi / f_lo) - 1), 256, 4096), which
    spreads the seeded batch over the whole [256, 4096] range."""
    MTSUN = 4.925491025873693e-06
    m1, m2, tref = p[0], p[1], p[11]
    Mc = (m1 * m2) ** 0.6 / (m1 + m2) ** 0.2 * MTSUN
    f_pn = (5.0 * Mc / np.maximum(tref, 1.0)) ** 0.375 / (8 * np.pi * Mc)
    f_lo = np.maximum(f_pn, 1.0 / t_data)
    f_hi = np.minimum(0.1 / ((m1 + m2) * MTSUN), 0.05)
    span = np.log10(np.maximum(f_hi / f_lo, 1.0 + 1e-12))
    return np.clip(np.round(2400 * (span - 1.0)), 256, 4096).astype(np.int32)


def bench_direct_ragged(args):
    """Uneven batching: direct likelihood of templates with per-template sparse
    grid lengths, fused, in one persistent launch sequence captured in a
    hipGraph, launch shape from the autotuner. Prints one JSON line."""
    import torch

    from bbhx_amd.autotune import tune_direct
    from bbhx_amd.cutils import timing_enable, timing_read
    from bbhx_amd.likelihood import Likelihood, fused_param_rows
    from bbhx_amd.utils.constants import YRSID_SI
    from bbhx_amd.utils.orbits import EqualArmlengthOrbits
    from bbhx_amd.utils.sensitivity import AET1SensitivityMatrix
    from bbhx_amd.waveformbuild import BBHWaveformFD
    from tests.params import MODES, draw

    T = args.direct_years * YRSID_SI
    n = int(T // args.direct_cadence)
    f = np.fft.rfftfreq(n, args.direct_cadence)[1:]
    if args.direct_bins:
        f = f[:args.direct_bins]
    if args.uniform_length:
        p = draw(args.batch, 1005)
        lengths = np.full(p.shape[1], args.uniform_length, dtype=np.int32)
    else:
        p = draw(args.batch, 1002, mtot=(1e4, 1e8))
        lengths = ragged_lengths(p, YRSID_SI)
    log(f"uneven batch: {p.shape[1]} templates, L in [{lengths.min()}, {lengths.max()}], sum {lengths.sum()}, "
        f"{f.size} dense bins")
    orbits = EqualArmlengthOrbits(dt=1e4, t_max=3.0 * YRSID_SI)
    wave = BBHWaveformFD(response_kwargs=dict(orbits=orbits, tdi2=True))
    sens = AET1SensitivityMatrix(f)
    psd = np.array([sens[0], sens[1], sens[2]])
    inj = wave(*p[:, 0], freqs=f, length=1024, modes=MODES, direct=False, fill=True)[0].cpu().numpy()
    rng = np.random.default_rng(1003)
    df = f[1] - f[0]
    data = inj + np.sqrt(psd / (4 * df)) * (rng.standard_normal(psd.shape) + 1j * rng.standard_normal(psd.shape))
    like = Likelihood(wave, f, data, psd)
    if args.launch:
        # a fixed launch shape (profiling: the trace then holds one shape only)
        c, bl, gr = (int(v) for v in args.launch.split(","))
        tuned = {"chunk": c, "block": bl, "grid": gr, "ms": None}
        log(f"fixed launch shape: chunk {c}, block {bl}, grid {gr}")
    else:
        log("autotuning the launch shape")
        tuned = tune_direct(like, p, lengths, MODES, reps=2,
                            progress=lambda c, b, g, ms: log(f"  chunk {c} block {b} grid {g}: {ms:.2f} ms"))
        log(f"autotuner: chunk {tuned['chunk']}, block {tuned['block']}, grid {tuned['grid']} ({tuned['ms']:.2f} ms)")
    plan = like.direct_plan(lengths, MODES, chunk=tuned["chunk"], block=tuned["block"], grid=tuned["grid"])
    rows, ts, te = fused_param_rows(like.waveform_gen, p, 0.0, 1.0, like.device)
    flags = torch.zeros(rows.shape[1], dtype=torch.int32, device=like.device)
    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(st):
        for _ in range(max(args.warmup, 1)):
            flags.zero_()
            like.get_ll_fused_device(plan, rows, ts, te, flags, MODES)
    torch.cuda.current_stream().wait_stream(st)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=st):
        flags.zero_()
        out = like.get_ll_fused_device(plan, rows, ts, te, flags, MODES)
    torch.cuda.synchronize()
    eager = like.get_ll_fused(p, lengths, modes=MODES, plan=plan)
    g.replay()
    torch.cuda.synchronize()
    diff = float(np.max(np.abs(out.cpu().numpy() - eager)))
    band = plan.band()
    dense = int(np.sum(np.maximum(band[:, 1] - band[:, 0], 0)))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.replay()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    # set-up / dense-band split from eager calls (the captured graph carries no timing events)
    timing_enable(True)
    for _ in range(2):
        flags.zero_()
        like.get_ll_fused_device(plan, rows, ts, te, flags, MODES)
    ms_setup, ms_eval, ncalls = timing_read()
    timing_enable(False)
    B = rows.shape[1]
    split = {}
    if ncalls:
        split = {"setup_ms": ms_setup / ncalls, "dense_ms": ms_eval / ncalls}
    log(f"graph replay: {ms:.3f} ms/step, {B / (ms * 1e-3):.4g} evals/s; |graph - eager| = {diff:.2e}; "
        f"plan {plan.nbytes() / 2**30:.2f} GiB; {dense} dense (bin, template) pairs "
        f"({dense / (ms * 1e-3):.4g}/s); split {split}")

    # FP32-bulk form of the same batch (phase, t(f) and inner products FP64), its own graph
    f32_info = None
    if not args.no_fp32:
        hh64 = like.h_h.real.cpu().numpy()
        flags.zero_()
        like.get_ll_fused_device(plan, rows, ts, te, flags, MODES, precision=1)  # allocates the records
        torch.cuda.synchronize()
        g32 = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g32, stream=st):
            flags.zero_()
            out32 = like.get_ll_fused_device(plan, rows, ts, te, flags, MODES)
        torch.cuda.synchronize()
        g32.replay()
        torch.cuda.synchronize()
        ll32 = out32.cpu().numpy()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g32.replay()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ms32 = (t1 - t0) * 1e3 / args.steps
        timing_enable(True)
        for _ in range(2):
            flags.zero_()
            like.get_ll_fused_device(plan, rows, ts, te, flags, MODES)
        s32, e32, n32 = timing_read()
        timing_enable(False)
        plan.set_precision(0)
        dll = np.abs(ll32 - eager)
        f32_info = {"ms": ms32, "evals_per_s": B / (ms32 * 1e-3), "speedup_vs_fp64": ms / ms32,
                    "setup_ms": s32 / max(n32, 1), "dense_ms": e32 / max(n32, 1),
                    "max_abs_dlogL_vs_fp64": float(dll.max()), "median_abs_dlogL_vs_fp64": float(np.median(dll)),
                    "max_rel_dlogL_vs_hh": float(np.max(dll / np.maximum(np.abs(hh64), 1.0))),
                    "templates_within_1e-3": int(np.sum(dll < 1e-3)),
                    "templates_in_fp64": int(np.sum(hh64 > 2.0e4))}
        log(f"FP32 bulk: {ms32:.3f} ms/step ({ms / ms32:.2f}x the FP64 form), dense {e32 / max(n32, 1):.2f} ms; "
            f"|dlogL| vs FP64 max {dll.max():.2e}, median {np.median(dll):.2e}; "
            f"{int(np.sum(dll < 1e-3))}/{B} within 1e-3")

    # the unfused inner product (the reference's direct_like kernel, a10) on a few
    # templates materialised by the kernel-by-kernel path: HBM-bound, 168 B per
    # dense bin and template (h read 48 and written back noise-weighted 48, as the
    # reference does, d 48, w 24); event-timed on the current stream
    unfused = None
    if not args.no_unfused:
        nb = min(8, B)
        Lu = int(lengths[0])
        tc, ist, iln = like.waveform_gen(*p[:, :nb], freqs=like.data_freqs, length=Lu, modes=MODES, direct=False,
                                         fill=False)
        tcs = [t.flatten() for t in tc]
        ptrs = np.asarray([t.data_ptr() for t in tcs], dtype=np.int64)
        dh_u = torch.zeros(nb, dtype=torch.complex128, device=like.device)
        hh_u = torch.zeros(nb, dtype=torch.complex128, device=like.device)
        args_u = (dh_u, hh_u, like.data_channels, like.noise_factors, ptrs, ist, iln, like.data_stream_length, nb)
        like.like_gen(*args_u)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 5
        e0.record()
        for _ in range(reps):
            like.like_gen(*args_u)
        e1.record()
        torch.cuda.synchronize()
        ums = e0.elapsed_time(e1) / reps
        ubytes = 168.0 * float(np.sum(iln))
        unfused = {"kernel": "k_direct_like", "templates": nb, "sparse_length": Lu, "dense_bins": int(np.sum(iln)),
                   "ms": ums, "bytes": ubytes, "GB_per_s": ubytes / (ums * 1e-3) / 1e9,
                   "hbm_frac": ubytes / (ums * 1e-3) / (HBM_PEAK_GBS * 1e9)}
        log(f"unfused inner product (k_direct_like): {nb} templates, {int(np.sum(iln))} dense bins, {ums:.3f} ms, "
            f"{unfused['GB_per_s']:.0f} GB/s ({unfused['hbm_frac']:.2f} of HBM peak)")
        # the template build of the same templates (the reference's InterpTDI, a8:
        # repack + k_interp_tdi), repeated on its own inputs: per dense bin f (8 B),
        # one segment index per mode (4 B each) and the 3 complex channels out
        # (48 B); per (template, mode, knot) the 36 coefficients and the knot once
        ir, spl = like.waveform_gen.interp_response, like.waveform_gen.spline
        tp = np.asarray([t.data_ptr() for t in ir.template_carrier], dtype=np.int64)
        iargs = (tp, like.data_freqs, *spl.container, ir.t_start, ir.t_end, ir.length, ir.data_length, ir.num_bin_all,
                 ir.num_modes, ir.ptrs, ir.start_inds, ir.lengths)
        like.waveform_gen.backend.InterpTDI_wrap(*iargs)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            like.waveform_gen.backend.InterpTDI_wrap(*iargs)
        e1.record()
        torch.cuda.synchronize()
        ims = e0.elapsed_time(e1) / reps
        ibytes = (56.0 + 4.0 * len(MODES)) * float(np.sum(iln)) + 296.0 * nb * len(MODES) * Lu
        unfused["template_build"] = {"kernels": "k_interp_pack + k_interp_tdi", "ms": ims, "bytes": ibytes,
                                     "GB_per_s": ibytes / (ims * 1e-3) / 1e9,
                                     "hbm_frac": ibytes / (ims * 1e-3) / (HBM_PEAK_GBS * 1e9)}
        log(f"unfused template build (InterpTDI): {nb} templates, {ims:.3f} ms, "
            f"{unfused['template_build']['GB_per_s']:.0f} GB/s ({unfused['template_build']['hbm_frac']:.2f} of HBM peak)")
        del ir, spl, iargs
        del tc, tcs, dh_u, hh_u, args_u

    # gradient (11 parameters) of the same batch: the adjoint launch sequence
    grad_info = None
    if not args.no_grad:
        for _ in range(max(args.warmup, 1)):
            flags.zero_()
            like.get_ll_grad_fused_device(plan, rows, ts, te, flags, MODES)
        torch.cuda.synchronize()
        timing_enable(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            flags.zero_()
            llg, _ = like.get_ll_grad_fused_device(plan, rows, ts, te, flags, MODES)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        gs, ge, gn = timing_read()
        timing_enable(False)
        gms = (t1 - t0) * 1e3 / args.steps
        same = bool(np.array_equal(llg.cpu().numpy(), eager))
        grad_info = {"grad_ms": gms, "grad_evals_per_s": B / (gms * 1e-3), "grad_cost_vs_forward": gms / ms,
                     "grad_setup_ms": gs / max(gn, 1), "grad_eval_ms": ge / max(gn, 1), "logL_equal_forward": same}
        log(f"gradient: {gms:.3f} ms/step ({gms / ms:.2f}x the forward), setup {gs / max(gn, 1):.2f} ms, adjoint "
            f"{ge / max(gn, 1):.2f} ms; logL equal to the forward's: {same}")

    # roofline of the dense-band kernel: the oracle's counted FP64 work per dense
    # (bin, template) x the pairs of this batch, over the kernel's event-timed duration
    model = json.loads((ROOT / "bbhx_amd" / "flop_model.json").read_text())["direct"]
    roofline = None
    if split:
        flops = model["flops_per_dense_bin_template"] * dense
        achieved = flops / (split["dense_ms"] * 1e-3) / 1e12
        # unique HBM bytes of the launch: data and noise weights of the bands once
        # (24 B per channel and bin, + 8 B frequency) and every segment record once
        unique = 80.0 * f.size + 320.0 * len(MODES) * float(lengths.sum())
        traffic = None
        tfile = ROOT / "profiles" / "k_dir_like.direct_ragged_1024.traffic.json"
        if tfile.exists() and B == 1024 and not args.uniform_length and args.direct_years == 1.0:
            traffic = json.loads(tfile.read_text())["hbm_bytes_per_launch"]
        roofline = {"bound": "fp64-valu", "compute": "FP64 VALU (no dense contraction, no MFMA; peak = FP64 vector peak)",
                    "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic, "kernel": "k_dir_like",
                    "kernel_ms": split["dense_ms"], "flops_per_launch": flops,
                    "flops_per_dense_bin_template": model["flops_per_dense_bin_template"],
                    "unique_hbm_bytes": unique,
                    "unique_hbm_frac": unique / (split["dense_ms"] * 1e-3) / (HBM_PEAK_GBS * 1e9)}

    cpu = None
    if not args.no_cpu:
        # the FP64 CPU oracle template by template at each template's own length
        # (its loops run over templates, so one template uses one core)
        from oracle.pipeline import OracleLikelihood, OracleWaveform

        log("timing the CPU oracle")
        owave = OracleWaveform(orbits.as_host_dict(), tdi2=True)
        olike = OracleLikelihood(owave, f, data, psd)
        t0 = time.perf_counter()
        done, worst = 0, 0.0
        while done < B and (time.perf_counter() - t0 < args.cpu_seconds or done == 0):
            llo = olike.get_ll(p[:, done:done + 1], length=int(lengths[done]), modes=MODES)[0][0]
            worst = max(worst, abs(llo - float(eager[done])))
            done += 1
        sec = time.perf_counter() - t0
        cpu = {"value": done / sec, "unit": "logL evals/s", "cores": 1, "kind": "port",
               "sample": f"FP64 CPU oracle, the first {done} templates of the batch one by one at their own "
                         f"sparse lengths, {sec:.1f} s; max |dlogL| vs GPU {worst:.2e}"}
        log(f"cpu oracle: {done / sec:.4g} evals/s on 1 core over {done} templates; max |dlogL| vs GPU {worst:.2e}")
    out = {
        "metric": "direct logL evals/sec, 1 GPU",
        "value": B / (ms * 1e-3),
        "unit": "logL evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded injection + Gaussian noise, analytic A/E/T PSD)",
        "config": {"workload": ("direct logL, uneven batch" if not args.uniform_length else "direct logL, large")
                   + ", fused, hipGraph replay", "templates": B, "data_years": args.direct_years,
                   "cadence_s": args.direct_cadence,
                   "lengths_min": int(lengths.min()), "lengths_max": int(lengths.max()),
                   "lengths_sum": int(lengths.sum()), "dense_bins": int(f.size), "modes": 6, "tdi": "AET tdi2",
                   "launch": {k: tuned[k] for k in ("chunk", "block", "grid")}, "max_abs_dlogL_graph_vs_eager": diff,
                   "dense_bin_templates": dense, **split, "fp32_bulk": f32_info, "gradient": grad_info,
                   "unfused_inner_product": unfused},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--nodes", type=int, default=128)
    ap.add_argument("--dense-bits", type=int, default=20)
    ap.add_argument("--cpu-sample", type=int, default=4096)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-grad", action="store_true", help="direct workloads: skip the gradient timing")
    ap.add_argument("--no-fp32", action="store_true", help="direct workloads: skip the FP32-bulk form")
    ap.add_argument("--no-unfused", action="store_true",
                    help="direct workloads: skip the unfused inner-product (k_direct_like) bandwidth measurement")
    ap.add_argument("--launch", default="", help="direct workloads: fixed 'chunk,block,grid' instead of autotuning")
    ap.add_argument("--no-grad-1024", action="store_true",
                    help="skip the gradient workload of 1024 templates at K = 256")
    ap.add_argument("--workload", default="heterodyned", choices=("heterodyned", "direct-ragged"))
    ap.add_argument("--direct-bins", type=int, default=0, help="truncate the dense data grid (testing)")
    ap.add_argument("--direct-years", type=float, default=1.0, help="direct workloads: data span [yr]")
    ap.add_argument("--direct-cadence", type=float, default=10.0, help="direct workloads: sampling cadence [s]")
    ap.add_argument("--uniform-length", type=int, default=0,
                    help="direct workloads: same sparse length for every template (large direct: 4096, with "
                         "--batch 512 --direct-years 4 --direct-cadence 5)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 256 MiB device-to-device copy (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the template batch shards by template "
              "with no inter-GPU exchange, so N GPUs would each run this on B/N templates", file=sys.stderr)
        sys.exit(2)

    import torch

    from bbhx_amd.cutils import timing_enable, timing_read

    if args.workload == "direct-ragged":
        if args.batch == 4096:
            args.batch = 8192
        bench_direct_ragged(args)
        return

    if args.copy_calibration:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 268435456 bytes read, 268435456 bytes written")
        del src, dst

    log("building problem")
    prob = build_problem(args.batch, args.nodes, args.dense_bits)
    het = prob["het"]
    rows, t_start, t_end = het.fused_inputs(prob["templates"], {})
    B = rows.shape[1]
    torch.cuda.synchronize()

    def step():
        return het.get_ll_device(rows, t_start, t_end)

    log("warmup")
    # the whole step (record set-up, residual/sum kernel, logL arithmetic) as one
    # hipGraph: the timed region replays it (no allocation, no host work per step)
    ms_per_step, ll_graph, graph = graph_time(step, args.steps, args.warmup)
    value = B / (ms_per_step * 1e-3)
    ll_eager = step().cpu().numpy()
    graph.replay()
    torch.cuda.synchronize()
    graph_diff = float(np.max(np.abs(ll_graph.cpu().numpy() - ll_eager)))
    ll_fused = ll_graph.cpu().numpy()
    ll_dev = ll_graph
    # set-up / evaluation split from eager calls (a captured graph carries no timing events)
    timing_enable(True)
    t2 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    ms_setup, ms_eval, ncalls = timing_read()
    timing_enable(False)
    ms_eager = (t3 - t2) * 1e3 / args.steps
    log(f"fused FP64 (hipGraph replay): {ms_per_step:.3f} ms/step, {value:.4g} evals/s; eager {ms_eager:.3f} ms/step; "
        f"setup {ms_setup / max(ncalls, 1):.3f} ms, eval {ms_eval / max(ncalls, 1):.3f} ms per call; "
        f"|graph - eager| = {graph_diff:.1e}")

    # gradient (11 parameters) of the same log-likelihoods: adjoint kernels, also
    # timed as hipGraph replays (the entry point forks a side stream and joins it
    # with events, which a graph captures)
    gms, (ll_g, grad), _ = graph_time(lambda: het.get_ll_grad_device(rows, t_start, t_end), args.steps, args.warmup)
    timing_enable(True)
    for _ in range(args.steps):
        het.get_ll_grad_device(rows, t_start, t_end)
    torch.cuda.synchronize()
    gms_setup, gms_eval, gn = timing_read()
    timing_enable(False)
    grad_rate = B / (gms * 1e-3)
    grad_ll_diff = float(torch.max(torch.abs(ll_g - ll_dev)).item())
    log(f"gradient (hipGraph replay): {gms:.3f} ms/step, {grad_rate:.4g} grad evals/s "
        f"({value / grad_rate:.2f}x the forward time); setup {gms_setup / max(gn, 1):.3f} ms, "
        f"eval {gms_eval / max(gn, 1):.3f} ms; max |logL(grad kernel) - logL(forward)| = {grad_ll_diff:.2e}")

    # gradient workload of 1024 templates at K = 256 sparse nodes (same reference and data model)
    grad_1024 = None
    if not args.no_grad_1024:
        log("gradient workload: 1024 templates, K = 256")
        prob_g = build_problem(1024, 256, args.dense_bits, seed=1003)
        het_g = prob_g["het"]
        rg, tsg, teg = het_g.fused_inputs(prob_g["templates"], {})
        fwd_ms, _, _ = graph_time(lambda: het_g.get_ll_device(rg, tsg, teg), args.steps, args.warmup)
        grd_ms, _, _ = graph_time(lambda: het_g.get_ll_grad_device(rg, tsg, teg), args.steps, args.warmup)
        grad_1024 = {"templates": 1024, "sparse_nodes": 256, "forward_ms": fwd_ms, "grad_ms": grd_ms,
                     "grad_evals_per_s": 1024 / (grd_ms * 1e-3), "grad_cost_vs_forward": grd_ms / fwd_ms}
        log(f"gradient (1024 x K=256): {grd_ms:.3f} ms/step, {grad_1024['grad_evals_per_s']:.4g} grad evals/s, "
            f"{grd_ms / fwd_ms:.2f}x the forward ({fwd_ms:.3f} ms)")
        del het_g, prob_g

    # the kernel-by-kernel FP64 path (the reference's structure) on the same templates
    t0 = time.perf_counter()
    n_plain = 3
    for _ in range(n_plain):
        ll_plain = het.get_ll(prob["templates"], fused=False)
    torch.cuda.synchronize()
    plain_rate = B / ((time.perf_counter() - t0) / n_plain)
    log(f"plain FP64 path: {plain_rate:.4g} evals/s; max |ll_fused - ll_plain| = "
        f"{np.max(np.abs(ll_fused - ll_plain)):.3e}")

    # algorithmic FP64 flops per launch of the fused evaluation kernel: the
    # oracle's per-node operation count per template (counted on this workload) x templates
    flop_model = json.loads((ROOT / "bbhx_amd" / "flop_model.json").read_text())
    flops = flop_model["eval_flops_per_template"] * B
    eval_s = (ms_eval / max(ncalls, 1)) * 1e-3
    achieved = flops / eval_s / 1e12 if eval_s > 0 else None
    traffic = None
    tfile = ROOT / "profiles" / "k_fused_eval.hdyn_4096x128.traffic.json"
    if tfile.exists() and B == 4096 and args.nodes == 128:
        traffic = json.loads(tfile.read_text())["hbm_bytes_per_launch"]
    roofline = {"bound": "fp64-valu", "compute": "FP64 VALU (no dense contraction, no MFMA; peak = FP64 vector peak)",
                "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": (achieved / FP64_PEAK_TFLOPS) if achieved else None, "traffic": traffic,
                "kernel": "k_fused_eval", "kernel_ms": ms_eval / max(ncalls, 1), "flops_per_launch": flops,
                "flop_rule": flop_model["counting_rules"]}

    # end to end: host parameters in, host logL out (parameter rows built and
    # uploaded, the fused launches, the logL copied back), as get_ll(params) runs
    het.get_ll(prob["templates"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n_e2e = max(3, args.steps // 4)
    for _ in range(n_e2e):
        het.get_ll(prob["templates"])
    e2e_ms = (time.perf_counter() - t0) * 1e3 / n_e2e
    log(f"end to end get_ll(params) (host in, host out): {e2e_ms:.3f} ms per call of {B} templates")

    cpu = None
    acc = None
    if not args.no_cpu:
        log("timing the CPU oracle")
        c = cpu_baseline(prob, min(args.cpu_sample, B), args.nodes, args.cpu_seconds)
        n = len(c["ll"])
        d = np.abs(ll_fused[:n] - c["ll"])
        floor = np.abs(c["ll_fma"] - c["ll"])
        acc = {"max_abs_dlogL_vs_cpu_oracle": float(d.max()),
               "max_rel_dlogL_vs_cpu_oracle": float(np.max(d / np.abs(c["ll"]))),
               "templates_within_1e-6": int(np.sum(d <= 1e-6)), "templates_checked": n,
               "oracle_fma_rounding_spread": float(floor.max()),
               "note": "oracle_fma_rounding_spread: max |dlogL| between the FP64 oracle and the same sources "
                       "compiled with FMA contraction (FP64 rounding alone)"}
        cpu = {"value": c["value"], "unit": "logL evals/s", "cores": c["cores"], "kind": "port",
               "sample": f"FP64 CPU oracle (OpenMP over templates), first {n} templates of the batch, "
                         f"{c['reps']} passes of {c['seconds_per_batch']:.3f} s",
               "value_1thread": c["value_1thread"]}
        log(f"cpu oracle: {c['value']:.4g} evals/s on {c['cores']} threads, {c['value_1thread']} on 1 thread; "
            f"max |dlogL| vs GPU = {d.max():.3e} ({int(np.sum(d <= 1e-6))}/{n} within 1e-6); oracle FMA "
            f"rounding spread {floor.max():.3e}")

    out = {
        "metric": "heterodyned logL evals/sec + grad evals/sec at 1 GPU, |dlogL|<1e-3 vs FP64",
        "value": value,
        "unit": "logL evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded injection + Gaussian noise, analytic A/E/T PSD)",
        "config": {"workload": "heterodyned logL, fused FP64, hipGraph replay", "templates": B,
                   "sparse_nodes": args.nodes, "eager_ms_per_step": ms_eager, "max_abs_dlogL_graph_vs_eager": graph_diff,
                   "dense_bins": 2**args.dense_bits, "modes": 6, "tdi": "AET tdi2",
                   "plain_fp64_path_evals_per_s": plain_rate, "end_to_end_get_ll_ms": e2e_ms,
                   "end_to_end_evals_per_s": B / (e2e_ms * 1e-3), "accuracy_vs_cpu_oracle": acc,
                   "grad_evals_per_s": grad_rate, "grad_cost_vs_forward": value / grad_rate,
                   "grad_params": "lnM, q, chi1, chi2, dist, phi_ref, cos inc, lam, sin beta, psi, t_ref",
                   "grad_1024x256": grad_1024},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
