"""The pGM dump of one evaluation (GROMACS GMX_PGM_DUMP): reader, writer, and the system a dump describes.

- A dump in the writer's layout, typed here line by line (comments, a covalent entry, the GVDW block), parses
  to the right fields.
- Round trip: a dump written by pgmhip.dump.write and read back is the same dump, number for number (17
  significant digits are exact for doubles).
- The reference's own reader: core/tests/run_dump.cpp, compiled from the reference sources into
  oracle/_ref/run_dump, parses a dump written here from an oracle evaluation and recomputes it with the
  reference core; its differences are at the solver tolerance.
- Dump.system() poses the oracle the same problem: recomputing reproduces the dumped terms, forces, dipoles
  and virial.
"""

import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from pgmhip import dump as D
from pgmhip import systems as S

RUN_DUMP = os.path.join(os.path.dirname(O.REF_LIB), "run_dump")

TEXT = """# pGM dump: n ncov beta rc K1 K2 K3 order ke
3 1 3.5 0.9 16 18 20 6 138.935458
# terms [kJ/mol]: direct reciprocal self background polarization; iterations residual
-1.5 -2.5 -3.5 -4.5 -5.5 7 1.25e-11
# box rows [nm]
2 0 0
0.5 2.1 0
-0.25 0.5 2.2
# virial (Xi = S^T/2) [kJ/mol]
1 2 3
4 5 6
7 8 9
# atom (global order): x(3) [nm] q radius alpha, pGM force (3) [kJ/mol/nm], induced dipole (3) [e nm], ...
0.1 0.2 0.3 -0.8 0.075 0.001 1 2 3 4 5 6 7 8 9
0.11 0.2 0.3 0.4 0.055 0.0004 -1 -2 -3 -4 -5 -6 -7 -8 -9
0.1 0.21 0.3 0.4 0.055 0.0004 0.5 0.5 0.5 0 0 0 0 0 0
# covalent dipoles: atom partner strength [e nm] (0-based, global)
0 1 0.01
# GVDW: energy (pairs) energy (overrides) energy (tail) [kJ/mol]; virial (3x3); per atom force
10 20 30
1 0 0
0 1 0
0 0 1
1 1 1
2 2 2
3 3 3
"""


def test_reads_the_writer_layout(tmp_path):
    p = tmp_path / "typed.dump"
    p.write_text(TEXT)
    d = D.read(str(p))
    assert d.n == 3 and d.grid == (16, 18, 20) and d.order == 6
    assert (d.beta, d.cutoff, d.ke) == (3.5, 0.9, 138.935458)
    assert d.terms == {"direct": -1.5, "reciprocal": -2.5, "self": -3.5, "background": -4.5, "polarization": -5.5}
    assert (d.iterations, d.residual) == (7, 1.25e-11)
    assert np.array_equal(d.box, [[2, 0, 0], [0.5, 2.1, 0], [-0.25, 0.5, 2.2]])
    assert np.array_equal(d.virial, np.arange(1, 10).reshape(3, 3))
    assert np.array_equal(d.pos[1], [0.11, 0.2, 0.3]) and d.q[0] == -0.8 and d.radius[2] == 0.055
    assert d.alpha[0] == 0.001 and np.array_equal(d.forces[1], [-1, -2, -3])
    assert np.array_equal(d.induced[0], [4, 5, 6]) and np.array_equal(d.permanent[1], [-7, -8, -9])
    assert list(d.cov_atom) == [0] and list(d.cov_partner) == [1] and list(d.cov_strength) == [0.01]
    assert d.gvdw["energy_tail"] == 30 and np.array_equal(d.gvdw["virial"], np.eye(3))
    assert np.array_equal(d.gvdw["forces"][2], [3, 3, 3])
    # and writes it back to the same numbers
    q = tmp_path / "again.dump"
    D.write(str(q), d)
    e = D.read(str(q))
    for k in ("box", "virial", "pos", "forces", "induced", "permanent", "cov_strength"):
        assert np.array_equal(getattr(d, k), getattr(e, k)), k
    assert e.gvdw["energy_pairs"] == 10 and np.array_equal(e.gvdw["forces"], d.gvdw["forces"])


@pytest.fixture(scope="module")
def evaluated():
    """An oracle evaluation of 512 waters (tau 1e-10, no predictor) as a dump."""
    sy = S.config0()
    sy.settings.update(tol=1e-10, max_iter=500, peek=0.0, predictor=0)
    o = O.Oracle(sy)
    _, f = o.compute(sy.pos, sy.box)
    d = D.from_evaluation(sy, o.terms(), f, o.strain(), o.induced(), o.permanent(), o.iterations(), o.residual())
    return sy, d


def test_round_trip_is_exact(tmp_path, evaluated):
    _, d = evaluated
    p = tmp_path / "w512.dump"
    D.write(str(p), d)
    e = D.read(str(p))
    for k in ("box", "virial", "pos", "q", "radius", "alpha", "forces", "induced", "permanent", "cov_atom",
              "cov_partner", "cov_strength"):  # fmt: skip
        assert np.array_equal(getattr(d, k), getattr(e, k)), k
    assert d.terms == e.terms and (d.beta, d.cutoff, d.grid, d.order, d.ke) == (e.beta, e.cutoff, e.grid, e.order, e.ke)
    assert e.iterations == d.iterations


def test_dump_poses_the_same_problem(evaluated):
    sy, d = evaluated
    sy2 = d.system()
    o = O.Oracle(sy2)
    _, f = o.compute(sy2.pos, sy2.box)
    t = o.terms()
    for k in D.TERMS:
        assert abs(t[k] - d.terms[k]) <= 1e-12 * abs(d.terms[k]) + 1e-12, k
    assert np.abs(f - d.forces).max() < 1e-10 * np.abs(d.forces).max()
    assert np.abs(0.5 * o.strain().T - d.virial).max() < 1e-10 * np.abs(d.virial).max()
    assert np.abs(o.induced() - d.induced).max() < 1e-12 and np.abs(o.permanent() - d.permanent).max() < 1e-14


@pytest.mark.skipif(not os.path.exists(RUN_DUMP), reason="oracle/_ref/run_dump not built (reference sources absent)")
def test_reference_reader_recomputes_our_dump(tmp_path, evaluated):
    _, d = evaluated
    p = tmp_path / "w512.dump"
    D.write(str(p), d)
    out = subprocess.run([RUN_DUMP, str(p)], capture_output=True, text=True, timeout=300, check=True).stdout
    print(out)
    rel = float(re.search(r"^total .* rel\s+(\S+)", out, re.M).group(1))
    fmax, frms = map(float, re.search(r"forces max\|diff\| (\S+) \(rms force (\S+)\)", out).groups())
    mu = float(re.search(r"mu max\|diff\| (\S+);", out).group(1))
    perm = float(re.search(r"perm max\|diff\| (\S+)", out).group(1))
    vd, vmax = map(float, re.search(r"virial max\|diff\| (\S+) \(max\|Xi\| (\S+)\)", out).groups())
    assert abs(rel) < 1e-11
    assert fmax < 1e-8 * frms
    assert mu < 1e-10 and perm < 1e-14
    assert vd < 1e-9 * vmax
