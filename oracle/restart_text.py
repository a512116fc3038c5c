"""Restart text writers as the reference writes them (TEST INFRASTRUCTURE ONLY).

CCSD_Helper.py:1107-1125: a csv.writer with delimiter '\\t', quotechar '|',
QUOTE_MINIMAL and lineterminator '\\n' writes one row per element in C order,
[i, a, numpy.around(F[i][a], decimals=precs)] or
[i, j, a, b, numpy.around(F[i][j][a][b], decimals=precs)]; the csv module
prints the numpy float64 scalar with str().
"""
from __future__ import annotations

import csv

import numpy as np


def _writer(f):
    return csv.writer(f, delimiter="\t", quotechar="|", quoting=csv.QUOTE_MINIMAL, lineterminator="\n")


def write_2data(F, FileName, precs):
    """CCSD_Helper.py:1107-1114"""
    with open(FileName, "w") as f:
        w = _writer(f)
        for i in range(F.shape[0]):
            for a in range(F.shape[1]):
                w.writerow([i, a, np.around(F[i][a], decimals=precs)])


def write_4data(F, FileName, precs):
    """CCSD_Helper.py:1116-1125"""
    with open(FileName, "w") as f:
        w = _writer(f)
        for i in range(F.shape[0]):
            for j in range(F.shape[1]):
                for a in range(F.shape[2]):
                    for b in range(F.shape[3]):
                        w.writerow([i, j, a, b, np.around(F[i][j][a][b], decimals=precs)])
