"""Writes tests/golden/gvdw_box.npz from the reference's GVDW test input (run in a tree that has it).

Input (a data file of the reference's tests): tests/data/gvdw_box.txt, made by tests/gvdw_input.py from a
prmtop and a restart: line 1 "n"; 3 lines of reduced box vectors [nm]; n lines "x y z radius has_lj" [nm];
then the excluded pairs "i j" (0-based).  Expected outputs: pmemd-pgm's VDWAALS of this box at rc = 10 A,
vdwmeth = 0, for the two repulsion forms with pmemd-pgm's global parameters, as core/tests/test_gvdw.cpp:62-68
holds them (kcal/mol; A_rep [kcal/mol], C6 [kcal/mol A^6], B): Slater 87500.0 / 594.825035 / 4.52 ->
784.7282; Gaussian 422.0 / 594.825035 / 0.9453 -> 868.3699.  GVDW acts on the atoms with has_lj only.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tests/data"

with open(os.path.join(REF, "gvdw_box.txt")) as fh:
    rows = [ln.split() for ln in fh if ln.strip()]
n = int(rows[0][0])
box = np.array([[float(v) for v in r] for r in rows[1:4]])
atoms = np.array([[float(v) for v in r] for r in rows[4 : 4 + n]])
excl = np.array([[int(v) for v in r] for r in rows[4 + n :]], np.int32).reshape(-1, 2)
np.savez_compressed(
    os.path.join(HERE, "gvdw_box.npz"),
    box=box, pos=atoms[:, :3], radius=atoms[:, 3], has_lj=atoms[:, 4].astype(np.int8), excluded=excl,
    rc=np.float64(1.0),
    model_rep=np.array([1, 0], np.int32),  # Slater, Gaussian
    model_arep_kcal=np.array([87500.0, 422.0]), model_c6_kcal_a6=np.array([594.825035, 594.825035]),
    model_b=np.array([4.52, 0.9453]), pmemd_kcal=np.array([784.7282, 868.3699]),
)  # fmt: skip
print("wrote", os.path.join(HERE, "gvdw_box.npz"), n, "atoms,", len(excl), "excluded pairs")
