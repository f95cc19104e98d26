"""PhenomD (l=2, m=2, n=0) QNM table: final spin -> ringdown and damping
frequency in units of 1/M_final (the reference's ``waveforms/ringdownphenomd.py``).
The numbers live in ``qnm_phenomd.txt`` next to this file."""

from pathlib import Path

import numpy as np

_tab = np.loadtxt(Path(__file__).with_name("qnm_phenomd.txt"))

dspin = 0.002
QNMData_a = np.ascontiguousarray(_tab[:, 0])
QNMData_fring = np.ascontiguousarray(_tab[:, 1])
QNMData_fdamp = np.ascontiguousarray(_tab[:, 2])

del _tab
