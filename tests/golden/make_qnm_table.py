"""Regenerate ``qnm_phenomd.txt``, the PhenomD (l=2, m=2, n=0) quasi-normal-mode
table used for the PhenomD ringdown and damping frequencies.

The table (final spin a in [-1, 1] with step 0.002, Re and Im of the QNM
frequency over 2 pi in units of the final mass) is physics data and has to be
the same numbers as the reference's ``waveforms/ringdownphenomd.py:3-3020``;
this script reads that file as text and writes the three columns. It is run
once, by hand, where the reference is available:

    python tests/golden/make_qnm_table.py /root/reference/src/bbhx/waveforms/ringdownphenomd.py
"""

import re
import sys
from pathlib import Path

import numpy as np


def parse_table(path):
    text = Path(path).read_text()
    spin_step = float(re.search(r"dspin\s*=\s*([0-9.eE+-]+)", text).group(1))
    out = {}
    for name in ("QNMData_a", "QNMData_fring", "QNMData_fdamp"):
        block = re.search(name + r"\s*=\s*np\.array\(\s*\[(.*?)\]\s*\)", text, re.S).group(1)
        out[name] = np.array([float(tok) for tok in block.replace("\n", " ").split(",") if tok.strip()])
    return spin_step, out["QNMData_a"], out["QNMData_fring"], out["QNMData_fdamp"]


def main(src):
    spin_step, a, fring, fdamp = parse_table(src)
    assert a.shape == fring.shape == fdamp.shape == (1001,)
    dst = Path(__file__).resolve().parents[2] / "bbhx_amd" / "waveforms" / "qnm_phenomd.txt"
    header = f"PhenomD QNM table, dspin={spin_step!r}; columns: final spin, f_ring, f_damp (units of 1/M_final)"
    np.savetxt(dst, np.column_stack([a, fring, fdamp]), fmt="%.17g", header=header)
    print(f"wrote {dst} ({a.size} rows)")


if __name__ == "__main__":
    main(sys.argv[1])
