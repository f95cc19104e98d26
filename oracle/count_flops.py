"""Counts the algorithmic floating-point work of the heterodyned-likelihood
forward path and writes it to bbhx_amd/flop_model.json.

TEST INFRASTRUCTURE ONLY. Runs the oracle's scalar code on an operation-counting
number type (opcount.h / opcount.cpp, built as oracle/libopcount.so) for a
sample of templates of the heterodyned-likelihood workload: reference binary =
first draw of seed 1000, templates = its N(0, 1e-3 x prior range)
perturbations (seed 1001), K sparse nodes, 2^20 dense bins at df = 1/yr, six
modes, TDI-2 A/E/T. The chain counted per template is the one
HeterodynedLikelihood.get_ll runs (likelihood.py:550-648): ringdown
frequencies, PhenomHM setup and evaluation on the K nodes, time/phase shift,
TDI response, mode sum, r = h/h0, heterodyned sums.

Counting rules: add, sub, mul, div, sqrt = 1 flop each; each libm call
(sin, cos, exp, log, pow, cbrt, atan, atan2, hypot) = 1 flop plus a separate
per-function tally.

The per-template work splits into a part independent of the node count
(ringdown frequencies and the per-mode coefficient setup) and a part linear in
it; the two are separated by counting at K and at 2K nodes (same band, same
log spacing, so the same mix of frequency regions).

    python oracle/count_flops.py [--sample 64] [--nodes 128]
"""

from __future__ import annotations

import argparse
import ctypes
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

LIBM = ("sin", "cos", "exp", "log", "pow", "cbrt", "atan", "atan2", "hypot")
ARITH = ("add", "mul", "div", "sqrt")


def _lib():
    path = ROOT / "oracle" / "libopcount.so"
    if not path.exists():
        raise FileNotFoundError("build it first: make -C oracle")
    lib = ctypes.CDLL(str(path))
    lib.opc_op_names.restype = ctypes.c_char_p
    return lib


def build_workload(nodes, dense_bits, seed=1000):
    from bbhx_amd.utils.constants import YRSID_SI
    from bbhx_amd.utils.orbits import EqualArmlengthOrbits
    from bbhx_amd.utils.sensitivity import AET1SensitivityMatrix
    from oracle.pipeline import OracleHeterodyne, OracleWaveform
    from tests.params import MODES, draw, perturb

    ref = draw(16, seed)[:, 0]
    f = np.arange(1, 2**dense_bits + 1) / YRSID_SI
    orbit = EqualArmlengthOrbits(dt=1e4, t_max=3.0 * YRSID_SI).as_host_dict()
    wave = OracleWaveform(orbit, tdi2=True)
    psd_fn = lambda x: [AET1SensitivityMatrix(x)[i] for i in range(3)]  # noqa: E731
    inj = wave(*ref, freqs=f, length=1024, modes=MODES, direct=False, fill=True, squeeze=True)[0]
    het = OracleHeterodyne(wave, f, inj, ref, nodes, psd_fn, reference_length=8096, modes=MODES)
    return ref, orbit, wave, het, perturb


def count(K, sample, dense_bits):
    """Operation counts summed over `sample` templates at K nodes, and the
    largest relative logL difference of the counted run against the oracle."""
    from oracle import _phenomd_rd_args, qnm_table
    from oracle.pipeline import tSSBfromLframe

    ref, orbit, wave, het, perturb = build_workload(K, dense_bits)
    p = perturb(ref, sample, 1001, scale=1e-3)
    B = p.shape[1]
    rows = np.ascontiguousarray(np.stack([p[0], p[1], p[2], p[3], p[4], p[5], p[7], p[8], p[9], p[10], p[11]]))
    YR = 31558149.763545603
    t_start = np.ascontiguousarray(tSSBfromLframe(0.0, p[8], p[9], 0.0) * np.ones(B))
    t_end = np.ascontiguousarray(tSSBfromLframe(1.0 * YR, p[8], p[9], 0.0) * np.ones(B))
    tab = qnm_table()
    arrs = _phenomd_rd_args(tab)
    qnm = np.ascontiguousarray(np.concatenate(arrs))
    table_len = len(arrs[0])
    h0 = np.ascontiguousarray(het.h0_sparse[0]).astype(np.complex128)
    dc = np.ascontiguousarray(het.data_constants).astype(np.complex128)
    freqs = np.ascontiguousarray(het.freqs, dtype=np.float64)
    links = np.ascontiguousarray(orbit["links"], dtype=np.int32)
    n_arr = np.ascontiguousarray(orbit["n_arr"], dtype=np.float64)
    x_arr = np.ascontiguousarray(orbit["x_arr"], dtype=np.float64)

    lib = _lib()
    nops = lib.opc_num_ops()
    names = lib.opc_op_names().decode().split(",")
    counts = np.zeros(nops)
    ll = np.zeros(B)
    D = ctypes.c_double
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    lib.opc_hdyn_forward(P(counts), P(ll), P(rows), B, P(freqs), K, P(h0), P(dc), P(qnm), table_len, D(tab["dspin"]),
                         P(t_start), P(t_end), D(0.0), D(orbit["dt"]), int(orbit["N"]), P(n_arr), P(x_arr), P(links),
                         D(orbit["armlength"]), 1)

    # the counted run must be the oracle's own computation
    ll_oracle = het.get_ll(p)[0] + 0.5 * het.reference_d_d
    err = float(np.max(np.abs(ll - ll_oracle) / np.maximum(1.0, np.abs(ll_oracle))))
    if err > 1e-9:
        raise RuntimeError(f"counted run differs from the oracle: {err}")

    return counts, err, B, names


def count_direct(sample, cadence=10.0, years=1.0):
    """Operation counts of the dense evaluation and direct inner products
    (orc_interp_tdi + orc_direct_like, counted) for the first `sample`
    templates of the uneven-batching workload (seed 1002, total mass
    log-uniform 1e4-1e8 Msun, sparse lengths by the bench's rule, data over
    `years` at `cadence` s), each at its own sparse length, as
    Likelihood.get_ll runs them (likelihood.py:140-225). Returns the counts,
    the number of dense (bin, template) pairs evaluated and the largest
    relative difference of the counted d_h, h_h against the uncounted oracle."""
    import bench
    import oracle.pipeline as pl
    from bbhx_amd.utils.constants import YRSID_SI
    from bbhx_amd.utils.orbits import EqualArmlengthOrbits
    from bbhx_amd.utils.sensitivity import AET1SensitivityMatrix
    from oracle.pipeline import OracleLikelihood, OracleWaveform
    from tests.params import MODES, draw

    T = years * YRSID_SI
    f = np.fft.rfftfreq(int(T // cadence), cadence)[1:]
    p = draw(8192, 1002, mtot=(1e4, 1e8))[:, :sample]
    lengths = bench.ragged_lengths(p, YRSID_SI)
    owave = OracleWaveform(EqualArmlengthOrbits(dt=1e4, t_max=3.0 * YRSID_SI).as_host_dict(), tdi2=True)
    sens = AET1SensitivityMatrix(f)
    psd = np.array([sens[0], sens[1], sens[2]])
    rng = np.random.default_rng(1003)
    data = np.sqrt(psd / (4 * (f[1] - f[0]))) * (rng.standard_normal(psd.shape) + 1j * rng.standard_normal(psd.shape))
    olike = OracleLikelihood(owave, f, data, psd)

    lib = _lib()
    nops = lib.opc_num_ops()
    names = lib.opc_op_names().decode().split(",")
    total = np.zeros(nops)
    pairs = 0
    worst = 0.0
    captured = {}
    orig_interp, orig_direct = pl.interp_tdi, pl.direct_like

    def interp_hook(*a, **k):
        captured["interp"] = a
        return orig_interp(*a, **k)

    def direct_hook(data_channels, noise_factors, templates, inds_start, ind_lengths, data_length, **k):
        nonlocal total, pairs, worst
        out = orig_direct(data_channels, noise_factors, templates, inds_start, ind_lengths, data_length, **k)
        (data_freqs, freqs, y, c1, c2, c3, t_start, t_end, length, B, M, inds_list, i_st, i_len) = captured["interp"]
        f64 = lambda x: np.ascontiguousarray(x, dtype=np.float64)  # noqa: E731
        i_len = np.ascontiguousarray(i_len, dtype=np.int32)
        ioff = np.zeros(B, dtype=np.int64)
        ioff[1:] = np.cumsum(M * i_len.astype(np.int64))[:-1]
        inds = np.ascontiguousarray(np.concatenate(inds_list), dtype=np.int32)
        counts = np.zeros(nops)
        dh = np.zeros(B, dtype=np.complex128)
        hh = np.zeros(B, dtype=np.complex128)
        P = lambda x: x.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        args = [f64(data_freqs), f64(freqs), f64(y), f64(c1), f64(c2), f64(c3), f64(t_start), f64(t_end)]
        dch = np.ascontiguousarray(data_channels, dtype=np.complex128)
        nf = f64(noise_factors)
        lib.opc_interp_direct(P(counts), P(dh), P(hh), P(args[0]), int(data_length), P(args[1]), P(args[2]),
                              P(args[3]), P(args[4]), P(args[5]), P(args[6]), P(args[7]), int(length), int(B), int(M),
                              P(inds), P(ioff), P(np.ascontiguousarray(i_st, dtype=np.int32)), P(i_len), P(dch),
                              P(nf))
        d_ref, h_ref = out
        worst = max(worst, float(np.max(np.abs(dh - d_ref) / np.maximum(np.abs(d_ref), 1e-300))),
                    float(np.max(np.abs(hh - h_ref) / np.maximum(np.abs(h_ref), 1e-300))))
        total += counts
        pairs += int(i_len.sum())
        return out

    pl.interp_tdi, pl.direct_like = interp_hook, direct_hook
    try:
        for b in range(p.shape[1]):
            olike.get_ll(p[:, b:b + 1], length=int(lengths[b]), modes=MODES)
    finally:
        pl.interp_tdi, pl.direct_like = orig_interp, orig_direct
    return total, pairs, worst, names, lengths


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sample", type=int, default=64)
    ap.add_argument("--nodes", type=int, default=128)
    ap.add_argument("--dense-bits", type=int, default=20)
    ap.add_argument("--direct-sample", type=int, default=8)
    ap.add_argument("--out", default=str(ROOT / "bbhx_amd" / "flop_model.json"))
    args = ap.parse_args()

    res = {}
    for K in (args.nodes, 2 * args.nodes):
        res[K] = count(K, args.sample, args.dense_bits)
    K = args.nodes
    counts, err, B, names = res[K]
    counts2 = res[2 * K][0]
    per_t = {n: counts[i] / B for i, n in enumerate(names)}
    tot = lambda c: sum(c[i] for i, n in enumerate(names) if n in ARITH + LIBM) / B  # noqa: E731
    libm_of = lambda c: sum(c[i] for i, n in enumerate(names) if n in LIBM) / B  # noqa: E731
    total_k, total_2k = tot(counts), tot(counts2)
    setup = 2 * total_k - total_2k
    evalk = total_k - setup
    libm_setup = 2 * libm_of(counts) - libm_of(counts2)
    M = 6
    model = {
        "workload": f"heterodyned logL forward, {M} modes, K = {K} nodes, TDI-2 AET, "
                    f"sample of {B} templates of the 1e-3 perturbation batch",
        "counting_rules": "add/sub/mul/div/sqrt = 1 flop; each libm call = 1 flop plus a per-function tally",
        "flops_per_template": total_k,
        "setup_flops_per_template": setup,
        "eval_flops_per_template": evalk,
        "libm_calls_per_template": libm_of(counts),
        "eval_libm_calls_per_template": libm_of(counts) - libm_setup,
        "per_template_counts": per_t,
        "eval_flops_per_mode_node": evalk / (M * K),
        "modes": M,
        "nodes": K,
        "generated_by": "python oracle/count_flops.py",
        "max_rel_ll_diff_vs_oracle": max(err, res[2 * K][1]),
    }
    dcounts, pairs, dworst, dnames, dlens = count_direct(args.direct_sample)
    if dworst > 1e-9:
        raise RuntimeError(f"counted direct run differs from the oracle: {dworst}")
    dflops = sum(dcounts[i] for i, n in enumerate(dnames) if n in ARITH + LIBM)
    model["direct"] = {
        "workload": f"dense evaluation + direct inner products (interp_tdi + direct_like), first "
                    f"{args.direct_sample} templates of the uneven batch (seed 1002), 1 yr at 10 s, six modes, "
                    f"TDI-2 AET, sparse lengths {dlens.tolist()}",
        "flops_per_dense_bin_template": dflops / pairs,
        "libm_calls_per_dense_bin_template": sum(dcounts[i] for i, n in enumerate(dnames) if n in LIBM) / pairs,
        "dense_bin_templates_counted": pairs,
        "max_rel_dh_hh_diff_vs_oracle": dworst,
    }
    Path(args.out).write_text(json.dumps(model, indent=1) + "\n")
    print(json.dumps(model, indent=1))


if __name__ == "__main__":
    main()
