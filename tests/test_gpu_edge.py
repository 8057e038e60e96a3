"""Edge cases of the HIP engine against the FP64 oracle (the cases core/tests/test_core.cpp and the reference's
Python tests reach through their parameters: polarizable and non-polarizable atoms together, inert atoms,
systems of one and two molecules).

- 512 waters where every fifth atom has alpha = 0, a few alpha = 1e-10 (at most 1e-9 nm^3 counts as none,
  core/pgm_constants.h:9) and a few have neither charge nor polarizability: FP64 terms and forces equal the
  oracle's to round-off, the non-polarizable atoms carry no induced dipole, mixed within its bound.
- One water and two waters in a 2 nm box (the cell list's fewer-than-5-cells path, a nearly empty grid): FP64
  equal to the oracle.
- PME grids with factors 5 and 7 and odd, unequal dimensions (35^3, 28 x 42 x 49, 40 x 30 x 50; the other tests'
  grids are 2^a 3^b or 20, 30, 40): FP64 equal to the oracle on the same grid, mixed within its bound.
- A batch of a one-water box and an ordinary snapshot equals its single evaluations.
"""

import dataclasses

import numpy as np
import pytest

import oracle as O
from conftest import rel_rms
from pgmhip import engine as E
from pgmhip import systems as S

pytestmark = pytest.mark.gpu


def _tight(sy):
    sy.settings.update(tol=1e-10, max_iter=500, peek=0.0, predictor=0)
    return sy


def _against_oracle(sy, precision="double", e_tol=1e-11, f_tol=1e-9):
    o = O.Oracle(sy)
    eo, fo = o.compute(sy.pos, sy.box)
    eng = E.Engine.for_system(sy, precision=precision)
    t, f, _ = eng.compute(sy.pos, sy.box)
    assert abs(E.total_energy(t) / eo - 1) < e_tol, (E.total_energy(t), eo)
    assert rel_rms(f, fo) < f_tol
    return eng, o


def test_partly_polarizable_and_inert_atoms():
    sy = _tight(S.config0())
    sy.alpha = sy.alpha.copy()
    sy.q = sy.q.copy()
    sy.alpha[::5] = 0.0
    sy.alpha[1::97] = 1e-10  # below the threshold: not polarizable
    inert = np.arange(2, sy.n, 151)
    inert = inert[sy.alpha[inert] > 0][:4]
    # keep the molecules neutral: the inert atom's charge moves to its oxygen
    for a in inert:
        o = a - a % 3
        if o != a:
            sy.q[o] += sy.q[a]
            sy.q[a] = 0.0
            sy.alpha[a] = 0.0
    eng, o = _against_oracle(sy)
    mu = eng.induced_dipoles()
    none = sy.alpha <= 1e-9
    assert np.abs(mu[none]).max() == 0.0
    assert np.abs(mu - o.induced()).max() < 1e-9 * np.abs(o.induced()).max()
    m = _tight(dataclasses.replace(sy))
    m.settings.update(tol=1e-6)
    _against_oracle(m, "mixed", 1e-6, 5e-5)


@pytest.mark.parametrize("nmol", [1, 2])
def test_one_and_two_molecules(nmol):
    sy = _tight(S.water_box(nmol, 2.0, 2, 5, 0.05, f"w{nmol}", S.base_settings(grid=(20, 20, 20))))
    eng, o = _against_oracle(sy, e_tol=1e-10, f_tol=1e-8)
    assert eng.num_pairs() == o.num_pairs()


@pytest.mark.parametrize("grid", [(35, 35, 35), (28, 42, 49), (40, 30, 50)])
def test_grids_with_factors_5_and_7(grid):
    """The transforms' radix-5 and radix-7 stages and odd, unequal dimensions: FP64 equal to the oracle on the
    same grid, mixed within the force bound."""
    sy = _tight(S.config0())
    sy.settings.update(grid=grid)
    _against_oracle(sy)
    m = _tight(dataclasses.replace(sy))
    m.settings.update(tol=1e-6)
    _against_oracle(m, "mixed", 1e-6, 5e-5)


def test_batch_with_a_one_molecule_box():
    from test_gpu_batch import _batch

    tiny = _tight(S.water_box(1, 2.0, 2, 5, 0.05, "w1", S.base_settings(grid=(20, 20, 20))))
    snap = S.config3_snapshot(9)
    snap.settings.update(tol=1e-10, max_iter=500, peek=0.0, predictor=0)
    systems = [tiny, snap]
    eng, off, pos, boxes = _batch(systems)
    terms, f = eng.compute_batch(pos, boxes)
    for k, sy in enumerate(systems):
        single = E.Engine.for_system(dataclasses.replace(sy), precision="double")
        ts, fs, _ = single.compute(sy.pos, sy.box)
        assert abs(E.total_energy(terms[k]) / E.total_energy(ts) - 1) < 1e-10, k
        assert np.abs(f[off[k] : off[k + 1]] - fs).max() < 1e-8 * np.abs(fs).max(), k
