"""Writes tests/golden/water512_pgmjax.npz from the reference's own test data (run in a tree that has it).

Inputs (data files of the reference's tests):
  tests/data/water512/pgm3p25_512.prmtop  POL_GAUSS parameters (text)
  tests/data/ref_w512_{q,qp,qi,qpi}_novdw.npz  pGM-JAX FP64 results for the four levels (positions, box,
  energy, forces, induced dipoles, iterations)
The fixture keeps those arrays as they are, plus the parameters in engine units (read with
pgmhip.prmtop, which reads the sections openmm/python/pgm_prmtop.py:54-115 reads).  Settings of the fixture:
tests/openmm_tests.sh:18-22 and tests/ref_pgmjax.py:36-40 (cutoff 0.9, beta 3.5, 36^3, order 6, tau 1e-10,
no predictor, no peek, ke 138.935458).
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "pgm-md_amd"))
from pgmhip import prmtop  # noqa: E402

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tests/data"

par = prmtop.pgm_parameters(os.path.join(REF, "water512", "pgm3p25_512.prmtop"))
out = {k: v for k, v in par.items()}
for level in ("q", "qp", "qi", "qpi"):
    d = np.load(os.path.join(REF, f"ref_w512_{level}_novdw.npz"))
    if level == "q":
        out["pos"] = d["pos"]
        out["H"] = d["H"]
    else:
        assert np.array_equal(out["pos"], d["pos"]) and np.array_equal(out["H"], d["H"])
    for key in ("energy", "forces", "mu", "iterations", "term_elec"):
        out[f"{level}_{key}"] = d[key]
np.savez_compressed(os.path.join(HERE, "water512_pgmjax.npz"), **out)
print({k: np.shape(v) for k, v in out.items()})
