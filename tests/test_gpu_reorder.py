"""The same atoms under new caller indices, and new parameters for the same atoms: the pair list is kept
(renamed, not rebuilt), and the results are those of the unpermuted run.

Mirrors the reference's reorder test (core/cuda/tests/test_cuda.cpp:205-242: setAtoms with a permutation and
global indices, setCovalent with remapped indices, the list builds unchanged) and continues a keyed mu4
trajectory through the reordering, as an MD code does after a pair search (core/cuda/pgm_cuda.cu:2519-2533,
2586-2632).  The iteration counts are the oracle's with the same history keys.
"""

import numpy as np
import pytest

import oracle as O
from conftest import rel_rms
from pgmhip import engine as E
from pgmhip import systems as S

pytestmark = pytest.mark.gpu


def _perm(n: int) -> np.ndarray:
    p = (np.arange(n) * 7919 + 13) % n   # new local k holds old atom p[k]
    assert len(np.unique(p)) == n
    return p


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_reorder_keeps_list_and_results(precision):
    sy = S.config0()
    n = sy.n
    traj = S.rigid_moves(sy, 11, 0.002, 6)
    perm = _perm(n)
    new_of = np.empty(n, np.int64)
    new_of[perm] = np.arange(n)
    o = O.Oracle(sy)
    plain = E.Engine.for_system(sy, precision=precision)
    moved = E.Engine.for_system(sy, precision=precision)
    for step, pos in enumerate(traj):
        o.set_history_key(step)
        eo, _ = o.compute(pos, sy.box)
        tp, fp, _ = plain.compute(pos, sy.box, new_step=True)
        if step == 3:   # the reordering, between two MD steps
            builds = moved.list_builds()
            moved.set_atoms(sy.q[perm], sy.radius[perm], sy.alpha[perm], global_index=perm, n_global=n)
            moved.set_covalent(new_of[sy.cov_atom], new_of[sy.cov_partner], sy.cov_strength)
        if step >= 3:
            tm, fm, _ = moved.compute(pos[perm], sy.box, new_step=True)
            fm_back = np.empty_like(fm)
            fm_back[perm] = fm
            assert moved.list_builds() == builds, "the reordering rebuilt the list"
            mu_back = moved.induced_dipoles()[new_of]
            mu = plain.induced_dipoles()
            assert np.abs(mu_back - mu).max() <= (1e-12 if precision == "double" else 1e-6) * np.abs(mu).max()
        else:
            tm, fm_back, _ = moved.compute(pos, sy.box, new_step=True)
        assert plain.iterations() == o.iterations() and moved.iterations() == o.iterations()
        et, em = E.total_energy(tp), E.total_energy(tm)
        if precision == "double":
            assert abs(em - et) <= 1e-10 * abs(et)
            assert np.abs(fm_back - fp).max() <= 1e-9 * np.abs(fp).max()
            assert abs(et - eo) <= 1e-9 * abs(eo)
        else:
            assert abs(em - et) <= 1e-8 * abs(et)
            assert rel_rms(fm_back, fp) < 1e-6


@pytest.mark.parametrize("new_strengths", [False, True])
def test_new_parameters_keep_list(new_strengths):
    """set_atoms with new charges, radii and polarizabilities for the same atoms (a fitting step), and
    set_covalent with the same entries (new strengths: only those are uploaded): no list build, and the results
    of a fresh engine with those parameters."""
    sy = S.config0()
    eng = E.Engine.for_system(sy, precision="double")
    eng.compute(sy.pos, sy.box)
    builds = eng.list_builds()
    q2, r2, a2 = sy.q * 1.05, sy.radius * 0.97, sy.alpha * 1.1
    s2 = sy.cov_strength * np.linspace(0.9, 1.1, len(sy.cov_strength)) if new_strengths else sy.cov_strength
    eng.set_atoms(q2, r2, a2)
    eng.set_covalent(sy.cov_atom, sy.cov_partner, s2)
    eng.reset_history()   # the same guess as the fresh engine's
    t, f, _ = eng.compute(sy.pos, sy.box)
    assert eng.list_builds() == builds
    sy2 = S.System(sy.name, sy.pos, sy.box, q2, r2, a2, sy.cov_atom, sy.cov_partner, s2, sy.settings)
    fresh = E.Engine.for_system(sy2, precision="double")
    tf, ff, _ = fresh.compute(sy.pos, sy.box)
    assert abs(E.total_energy(t) / E.total_energy(tf) - 1) < 1e-12
    assert np.abs(f - ff).max() <= 1e-10 * np.abs(ff).max()
    o = O.Oracle(sy2)
    eo, fo = o.compute(sy.pos, sy.box)
    assert abs(E.total_energy(t) / eo - 1) < 1e-10
