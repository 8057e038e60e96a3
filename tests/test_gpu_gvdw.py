"""GVDW in the HIP engine's final pair pass (pgm_set_gvdw), against the oracle's brute-force GVDW sum (pinned to
pmemd-pgm's box energies, mpmath and the reference's pair function in tests/test_gvdw_oracle.py).

- The golden box (tests/golden/gvdw_box.npz: 1536 atoms of a truncated-octahedron cell, the reference's
  tests/data/gvdw_box.txt) at rc = rvdw = 1 nm with the excluded pairs skipped: the engine's GVDW energy equals
  pmemd-pgm's VDWAALS to 1e-3 kcal/mol (the reference's tolerance, core/tests/test_gvdw.cpp:86) and the oracle's
  to FP64 round-off; forces and strain to the oracle's.  Mixed precision (FP32 pair arithmetic, FP64 sums):
  energy within 1e-5 relative, forces within the 5e-5 relative-RMS force bound of the mixed electrostatics.
- Water with charges, polarizabilities, covalent dipoles and GVDW, the intramolecular pairs skipped: the
  engine's terms, forces and strain are the oracle's electrostatics plus the oracle's GVDW sum.
- rvdw = 0 turns GVDW off again; rvdw beyond the cutoff and backward with GVDW on are refused.
"""

import os

import numpy as np
import pytest

import oracle as O
from pgmhip import engine as E
from pgmhip import systems as S

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KCAL = 4.184


def golden():
    return np.load(os.path.join(HERE, "golden", "gvdw_box.npz"))


def params(g, k):
    lj = g["has_lj"].astype(float)
    A = g["model_arep_kcal"][k] * KCAL
    C6 = g["model_c6_kcal_a6"][k] * KCAL * 1e-6
    return np.sqrt(A) * lj, np.sqrt(C6) * lj, np.full(len(lj), g["model_b"][k])


def box_engine(g, precision):
    n = len(g["pos"])
    eng = E.Engine(precision=precision, cutoff=1.0, beta=3.0, grid=(32, 32, 32), order=6, tol=1e-8, max_iter=100,
                   peek=0.0, predictor=0, polarizable=0)  # fmt: skip
    eng.set_atoms(np.zeros(n), g["radius"], np.zeros(n))
    eng.set_covalent([], [], [])
    return eng


@pytest.mark.parametrize("k", [0, 1])
def test_golden_box_fp64(k):
    g = golden()
    sa, sc, b = params(g, k)
    rep, rc = int(g["model_rep"][k]), float(g["rc"])
    ex = g["excluded"]
    eng = box_engine(g, "double")
    eng.set_gvdw(rep, rc, sa, sc, b, ex[:, 0], ex[:, 1])
    t, f, S3 = eng.compute(g["pos"], g["box"], strain=True)
    Eo, Fo, So = O.gvdw_sum(g["pos"], g["box"], g["radius"], sa, sc, b, rep, rc, ex[:, 0], ex[:, 1])
    print(f"model {k}: engine {t['gvdw'] / KCAL:.6f} kcal/mol, oracle {Eo / KCAL:.6f}, pmemd-pgm {g['pmemd_kcal'][k]}")
    assert abs(t["gvdw"] / KCAL - g["pmemd_kcal"][k]) < 1e-3
    assert abs(t["gvdw"] / Eo - 1) < 1e-12
    assert np.abs(f - Fo).max() < 1e-10 * np.abs(Fo).max()
    assert np.abs(S3 - So).max() < 1e-10 * np.abs(So).max()
    assert abs(t["direct"]) + abs(t["reciprocal"]) < 1e-12  # no charges


@pytest.mark.parametrize("k", [0, 1])
def test_golden_box_mixed(k):
    g = golden()
    sa, sc, b = params(g, k)
    rep, rc = int(g["model_rep"][k]), float(g["rc"])
    ex = g["excluded"]
    eng = box_engine(g, "mixed")
    eng.set_gvdw(rep, rc, sa, sc, b, ex[:, 0], ex[:, 1])
    t, f, _ = eng.compute(g["pos"], g["box"])
    Eo, Fo, _ = O.gvdw_sum(g["pos"], g["box"], g["radius"], sa, sc, b, rep, rc, ex[:, 0], ex[:, 1])
    rel = abs(t["gvdw"] / Eo - 1)
    print(f"model {k} mixed: energy {rel:.2e} relative, forces {E.rel_rms(f, Fo):.2e} rel-RMS")
    assert rel < 1e-5
    assert E.rel_rms(f, Fo) < 5e-5


def _water_with_gvdw():
    sy = S.config0()
    sy.settings.update(tol=1e-10, max_iter=500, peek=0.0, predictor=0)
    n = sy.n
    o = np.arange(0, n, 3)
    sa = np.zeros(n)
    sc = np.zeros(n)
    sa[o] = np.sqrt(422.0 * KCAL)  # Gaussian repulsion on O
    sc[o] = np.sqrt(594.825035 * KCAL * 1e-6)
    sc[o + 1] = sc[o + 2] = 0.3 * sc[o]  # a little dispersion on H
    b = np.full(n, 0.9453)
    si = np.concatenate([o, o, o + 1])
    sj = np.concatenate([o + 1, o + 2, o + 2])
    return sy, sa, sc, b, si, sj


def test_water_electrostatics_plus_gvdw():
    sy, sa, sc, b, si, sj = _water_with_gvdw()
    rvdw = 0.8
    eng = E.Engine.for_system(sy, precision="double")
    eng.set_gvdw(0, rvdw, sa, sc, b, si, sj)
    t, f, S3 = eng.compute(sy.pos, sy.box, strain=True)
    o = O.Oracle(sy)
    _, fo = o.compute(sy.pos, sy.box)
    Eg, Fg, Sg = O.gvdw_sum(sy.pos, sy.box, sy.radius, sa, sc, b, 0, rvdw, si, sj)
    to = o.terms()
    for key in ("direct", "reciprocal", "self", "background", "polarization"):
        assert abs(t[key] - to[key]) <= 1e-10 * abs(to[key]) + 1e-9, key
    assert abs(t["gvdw"] / Eg - 1) < 1e-12
    ftot = fo + Fg
    assert np.abs(f - ftot).max() < 1e-9 * np.abs(ftot).max()
    Stot = o.strain() + Sg
    assert np.abs(S3 - Stot).max() < 1e-9 * np.abs(Stot).max()
    # off again
    eng.set_gvdw(0, 0.0, sa, sc, b)
    t2, f2, _ = eng.compute(sy.pos, sy.box)
    assert t2["gvdw"] == 0.0
    assert np.abs(f2 - fo).max() < 1e-9 * np.abs(fo).max()


def test_gvdw_errors():
    sy, sa, sc, b, si, sj = _water_with_gvdw()
    eng = E.Engine.for_system(sy, precision="double")
    with pytest.raises(E.PgmError, match="rvdw"):
        eng.set_gvdw(0, sy.settings["cutoff"] + 0.1, sa, sc, b)
    with pytest.raises(E.PgmError, match="repulsion"):
        eng.set_gvdw(2, 0.5, sa, sc, b)
    with pytest.raises(E.PgmError, match="out of range"):
        eng.set_gvdw(0, 0.5, sa, sc, b, [0], [sy.n])
    eng.set_gvdw(0, 0.5, sa, sc, b, si, sj)
    eng.compute(sy.pos, sy.box)
    with pytest.raises(E.PgmError, match="GVDW"):
        eng.backward(sy.pos, sy.box, np.zeros((sy.n, 3)), 1.0)


def test_batch_with_gvdw_matches_single_systems():
    """GVDW in a batch (per-atom parameters and skipped pairs concatenated with the atom offsets): every system's
    GVDW energy, forces and other terms equal its single-system evaluation."""
    import dataclasses

    systems = [S.config3_snapshot(11), S.config3_snapshot(23), S.config3_snapshot(37)]
    for sy in systems:
        sy.settings.update(tol=1e-12, max_iter=2000, predictor=0, peek=0.0)
    rvdw = 0.8

    def gv(sy):
        n = sy.n
        o = np.arange(0, n, 3)
        sa, sc = np.zeros(n), np.zeros(n)
        sa[o] = np.sqrt(422.0 * KCAL)
        sc[o] = np.sqrt(594.825035 * KCAL * 1e-6)
        sc[o + 1] = sc[o + 2] = 0.3 * sc[o]
        return sa, sc, np.full(n, 0.9453), np.concatenate([o, o, o + 1]), np.concatenate([o + 1, o + 2, o + 2])

    off = np.concatenate([[0], np.cumsum([s.n for s in systems])]).astype(np.int32)
    cat = lambda k: np.concatenate([getattr(s, k) for s in systems])  # noqa: E731
    st = systems[0].settings
    keys = ("cutoff", "beta", "order", "tol", "max_iter", "peek", "predictor", "ke", "skin", "polarizable")
    eng = E.Engine(precision="double", grid=tuple(st["grid"]), **{k: st[k] for k in keys})
    eng.set_atoms(cat("q"), cat("radius"), cat("alpha"))
    eng.set_covalent(np.concatenate([s.cov_atom + off[i] for i, s in enumerate(systems)]),
                     np.concatenate([s.cov_partner + off[i] for i, s in enumerate(systems)]), cat("cov_strength"))
    eng.set_batch(off, [s.settings["grid"] for s in systems])
    parts = [gv(s) for s in systems]
    eng.set_gvdw(0, rvdw, *[np.concatenate([p[k] for p in parts]) for k in range(3)],
                 np.concatenate([p[3] + off[i] for i, p in enumerate(parts)]),
                 np.concatenate([p[4] + off[i] for i, p in enumerate(parts)]))  # fmt: skip
    terms, f = eng.compute_batch(np.concatenate([s.pos for s in systems]), np.stack([s.box for s in systems]))
    for k, sy in enumerate(systems):
        single = E.Engine.for_system(dataclasses.replace(sy), precision="double")
        single.set_gvdw(0, rvdw, *parts[k])
        ts, fs, _ = single.compute(sy.pos, sy.box)
        assert abs(terms[k]["gvdw"] / ts["gvdw"] - 1) < 1e-11, k
        assert abs(E.total_energy(terms[k]) / E.total_energy(ts) - 1) < 1e-11, k
        fk = f[off[k] : off[k + 1]]
        assert np.abs(fk - fs).max() < 1e-8 * np.abs(fs).max(), k
