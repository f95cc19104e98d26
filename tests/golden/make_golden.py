"""Write the golden vectors of the reference's own CPU code (oracle/_ref).

Run where the reference sources are available (they are compiled by
``make -C oracle ref`` from the reference tree, not copied):

    python tests/golden/make_golden.py

Writes ``tests/golden/reference_kats.npz``: inputs and outputs of
PhenomHMWaveform.cu (HM with 6 modes, PhenomD, an equal-mass case), both
ringdown-frequency functions, WaveformBuild.cu's direct_sum and InterpTDI and
Likelihood.cu's hdyn (CPU loop) and direct_like, on seeded inputs.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402

MSUN_SI = 1.98848e+30
MTSUN_SI = 4.925491025873693e-06
PC_SI = 3.0856775814913674e+16


def phm_inputs(rng, B):
    M = np.exp(rng.uniform(np.log(1e4), np.log(1e8), B))
    q = rng.uniform(1, 10, B)
    m1, m2 = M * q / (1 + q), M / (1 + q)
    c1 = rng.uniform(-0.99, 0.99, B)
    c2 = rng.uniform(-0.99, 0.99, B)
    m1[0] = m2[0]  # equal masses
    c1[1], c2[1] = 0.0, 0.0
    c1[2], c2[2] = 0.99, 0.99
    c1[3], c2[3] = -0.99, -0.99
    dist = rng.uniform(10, 100, B) * 1e9 * PC_SI
    return m1, m2, c1, c2, dist


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libbbhx_ref.so is not built (make -C oracle ref)")
    rng = np.random.default_rng(20240611)
    out = {}

    # ---- PhenomHM, 6 modes, 40 log-spaced frequencies per mode over Mf in [1e-4, 0.6] m/2
    B, L = 12, 40
    ells = np.array([2, 3, 4, 2, 3, 4], np.int32)
    mms = np.array([2, 3, 4, 1, 2, 3], np.int32)
    m1, m2, c1, c2, dist = phm_inputs(rng, B)
    frd, fdm = oracle.get_phenomd_ringdown_frequencies(m1, m2, c1, c2, use_ref=True)
    hr, hd = oracle.get_phenomhm_ringdown_frequencies(m1, m2, c1, c2, ells, mms, use_ref=True)
    Mf_RD = np.concatenate([hr, frd[:, None]], 1).ravel()
    Mf_DM = np.concatenate([hd, fdm[:, None]], 1).ravel()
    Msec = (m1 + m2) * MTSUN_SI
    base = np.logspace(-4, np.log10(0.6), L)
    freqs = (base[None, None, :] / Msec[:, None, None] * mms[None, :, None] / 2).ravel()
    wf = np.zeros(3 * B * 6 * L)
    oracle.waveform_amp_phase(wf, ells, mms, freqs, m1 * MSUN_SI, m2 * MSUN_SI, c1, c2, dist, np.zeros(B), 6, L, B,
                              Mf_RD, Mf_DM, 0, use_ref=True)
    out.update(phm_m1=m1, phm_m2=m2, phm_chi1=c1, phm_chi2=c2, phm_dist=dist, phm_ells=ells, phm_mms=mms,
               phm_freqs=freqs, phm_Mf_RD=Mf_RD, phm_Mf_DM=Mf_DM, phm_out=wf, phm_frd=frd, phm_fdm=fdm, phm_hr=hr,
               phm_hd=hd)

    # ---- PhenomD
    freqsD = (base[None, :] / Msec[:, None]).ravel()
    wfD = np.zeros(3 * B * L)
    oracle.waveform_amp_phase(wfD, np.array([2], np.int32), np.array([2], np.int32), freqsD, m1 * MSUN_SI,
                              m2 * MSUN_SI, c1, c2, dist, np.zeros(B), 1, L, B, frd, fdm, 1, use_ref=True)
    out.update(phd_freqs=freqsD, phd_out=wfD)

    # ---- direct_sum
    Bs, Ms, Ls = 3, 6, 40
    buf = rng.standard_normal(9 * Bs * Ms * Ls)
    n = Bs * Ms * Ls
    buf[:n] = np.abs(buf[:n])
    buf[n:2 * n] *= 1e4
    buf[2 * n:3 * n] = rng.uniform(-1e6, 3e7, n)
    ts = np.array([0.0, 5e5, 1e6])
    te = np.array([2e7, 0.0, 2.5e7])
    ds = oracle.direct_sum(buf, Bs, Ls, Ms, ts, te, use_ref=True)
    out.update(ds_buf=buf, ds_ts=ts, ds_te=te, ds_out=ds)

    # ---- InterpTDI
    Bi, Mi, Li, N = 3, 2, 30, 500
    df_ = np.arange(1, N + 1) * 1e-5
    x = np.sort(rng.uniform(2e-4, 4.8e-3, (Bi, Mi, Li)), axis=-1)
    y = rng.standard_normal(9 * Bi * Mi * Li)
    y[2 * Bi * Mi * Li:3 * Bi * Mi * Li] = rng.uniform(0, 1e7, Bi * Mi * Li)
    cc1, cc2, cc3 = (rng.standard_normal(9 * Bi * Mi * Li) for _ in range(3))
    st = np.searchsorted(df_, x.min(axis=(1, 2)), side="right").astype(np.int32)
    et = np.searchsorted(df_, x.max(axis=(1, 2)), side="right").astype(np.int32)
    inds = [np.concatenate([np.searchsorted(x[i, j], df_[st[i]:et[i]], side="right") - 1 for j in range(Mi)]).astype(
        np.int32) for i in range(Bi)]
    ln = (et - st).astype(np.int32)
    its = np.zeros(Bi)
    ite = np.full(Bi, 5e6)
    tm = oracle.interp_tdi(df_, x.ravel(), y, cc1, cc2, cc3, its, ite, Li, Bi, Mi, inds, st, ln, use_ref=True)
    out.update(it_df=df_, it_x=x, it_y=y, it_c1=cc1, it_c2=cc2, it_c3=cc3, it_st=st, it_len=ln,
               it_inds=np.concatenate(inds), it_ts=its, it_te=ite, it_out=np.concatenate([t.ravel() for t in tm]))

    # ---- hdyn (reference CPU loop)
    Bh, K = 5, 50
    fK = np.sort(rng.uniform(1e-4, 1e-2, K))
    r = (1 + 0.01 * rng.standard_normal((K, 3, Bh))) * np.exp(1j * 0.1 * rng.standard_normal((K, 3, Bh)))
    dc = (rng.standard_normal(12 * K) + 1j * rng.standard_normal(12 * K)) * 1e3
    o1, o2 = oracle.hdyn(r.ravel(), dc, fK, Bh, K, use_ref=True)
    out.update(hd_f=fK, hd_r=r.ravel(), hd_dc=dc, hd_dh=o1, hd_hh=o2)

    # ---- direct_like
    dat = rng.standard_normal((3, N)) + 1j * rng.standard_normal((3, N))
    w = rng.uniform(0.5, 2.0, 3 * N)
    tmpl = [t.copy() for t in tm]
    dh, hh = oracle.direct_like(dat.ravel(), w, tmpl, st, ln, N, use_ref=True)
    out.update(dl_data=dat.ravel(), dl_w=w, dl_dh=dh, dl_hh=hh)

    dst = Path(__file__).with_name("reference_kats.npz")
    np.savez_compressed(dst, **out)
    print(f"wrote {dst} ({dst.stat().st_size / 1024:.1f} KiB)")


if __name__ == "__main__":
    main()
