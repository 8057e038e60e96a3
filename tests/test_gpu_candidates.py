"""The caller's pair list (pgm_set_candidates, Engine::setCandidates of core/cuda/pgm_cuda.h:68-73) in the HIP
engine.

Candidates: every pair within rc + 0.1 nm of 512 waters (so some are beyond the cutoff and skipped), in segments
of one atom i and one periodic image t, x = r_i + t.box - r_j, as a pair search with shifts (GROMACS's)
produces them.

- FP64: the oracle's energies and forces (its own minimum-image pairs) at tau 1e-10; mixed within the 5e-5
  force bound.
- The images are applied, not recomputed: atoms moved by whole box vectors, with the images adjusted, give the
  same results (positions are used as given); a pair listed at a far image (beyond the cutoff) counts as
  absent: the same results as a list without it, and not those of the full list.
- The adjoint over a caller's list equals the engine's own; pgm_set_atoms returns to the engine's own list.
- Refused: a batch, GVDW, images beyond 7 box vectors, indices out of range, an atom paired with itself.
"""

import numpy as np
import pytest
from scipy.spatial import cKDTree

import oracle as O
from conftest import rel_rms
from pgmhip import engine as E
from pgmhip import systems as S

pytestmark = pytest.mark.gpu


def _water(tol=1e-10):
    sy = S.config0()
    sy.settings.update(tol=tol, max_iter=500, peek=0.0, predictor=0)
    return sy


def _pairs(pos, box, r):
    """Pairs i < j within r (cubic box) with the image t of the minimum-image displacement of pos as given."""
    L = box[0, 0]
    w = pos - L * np.floor(pos / L)
    w[w >= L] -= L  # rounding can land exactly on L
    pr = cKDTree(w, boxsize=L).query_pairs(r, output_type="ndarray")
    i, j = pr[:, 0], pr[:, 1]
    t = -np.rint((pos[i] - pos[j]) / L).astype(np.int32)
    return i.astype(np.int32), j.astype(np.int32), t


def _segments(i, j, t):
    """Segments of one (i, t): seg_i, seg_shift, seg_begin, cand_j."""
    code = (t[:, 0] + 8) + 16 * (t[:, 1] + 8) + 256 * (t[:, 2] + 8)
    o = np.lexsort((j, code, i))
    i, j, t, code = i[o], j[o], t[o], code[o]
    key = i.astype(np.int64) * 4096 + code
    starts = np.concatenate([[0], np.nonzero(np.diff(key))[0] + 1])
    seg_begin = np.concatenate([starts, [len(i)]]).astype(np.int32)
    return i[starts], t[starts], seg_begin, j


def _engine(sy, precision, segs=None):
    eng = E.Engine.for_system(sy, precision=precision)
    if segs is not None:
        eng.set_candidates(*segs)
    return eng


def test_fp64_and_mixed_against_the_oracle():
    sy = _water()
    o = O.Oracle(sy)
    eo, fo = o.compute(sy.pos, sy.box)
    segs = _segments(*_pairs(sy.pos, sy.box, sy.settings["cutoff"] + 0.1))
    d = _engine(sy, "double", segs)
    td, fd, _ = d.compute(sy.pos, sy.box)
    assert d.num_candidates() == 2 * len(segs[3])
    assert d.num_pairs() == o.num_pairs()
    assert d.iterations() == o.iterations()
    assert abs(E.total_energy(td) / eo - 1) < 1e-11
    assert rel_rms(fd, fo) < 1e-10
    m = _engine(_water(1e-6), "mixed", segs)
    tm, fm, _ = m.compute(sy.pos, sy.box)
    assert rel_rms(fm, fo) < 5e-5
    assert abs(E.total_energy(tm) / eo - 1) < 1e-6


def test_images_are_applied_as_given():
    sy = _water()
    ref = _engine(sy, "double", _segments(*_pairs(sy.pos, sy.box, sy.settings["cutoff"] + 0.1)))
    t0, f0, _ = ref.compute(sy.pos, sy.box)
    # every atom moved by whole box vectors; the images follow
    k = np.random.default_rng(3).integers(-2, 3, (sy.n, 3))
    pos2 = sy.pos + k * sy.box[0, 0]
    eng = _engine(sy, "double", _segments(*_pairs(pos2, sy.box, sy.settings["cutoff"] + 0.1)))
    t2, f2, _ = eng.compute(pos2, sy.box)
    assert abs(E.total_energy(t2) / E.total_energy(t0) - 1) < 1e-11
    assert rel_rms(f2, f0) < 1e-10
    # one pair at a far image (1 box vector off, beyond the cutoff) is absent, not the minimum image
    i, j, t = _pairs(sy.pos, sy.box, sy.settings["cutoff"] + 0.1)
    r = np.linalg.norm(sy.pos[i] - sy.pos[j] + t * sy.box[0, 0], axis=1)
    p = int(np.argmin(r))  # the closest pair: its removal changes the energy visibly
    far = t.copy()
    far[p, 0] += 1
    keep = np.arange(len(i)) != p
    a = _engine(sy, "double", _segments(i, j, far))
    b = _engine(sy, "double", _segments(i[keep], j[keep], t[keep]))
    ta, fa, _ = a.compute(sy.pos, sy.box)
    tb, fb, _ = b.compute(sy.pos, sy.box)
    assert a.num_pairs() == b.num_pairs() == ref.num_pairs() - 1
    assert abs(E.total_energy(ta) / E.total_energy(tb) - 1) < 1e-13
    assert rel_rms(fa, fb) < 1e-12
    assert rel_rms(fa, f0) > 1e-6


def test_adjoint_and_back_to_the_own_list():
    sy = _water()
    g = np.random.default_rng(5).normal(0.0, 1e-3, (sy.n, 3))
    own = _engine(sy, "double")
    t0, f0, _ = own.compute(sy.pos, sy.box)
    g0 = own.backward(sy.pos, sy.box, g, 0.5)
    ext = _engine(sy, "double", _segments(*_pairs(sy.pos, sy.box, sy.settings["cutoff"] + 0.05)))
    ext.compute(sy.pos, sy.box)
    g1 = ext.backward(sy.pos, sy.box, g, 0.5)
    for key in g0:
        assert np.abs(g1[key] - g0[key]).max() <= 1e-9 * np.abs(g0[key]).max(), key
    builds = ext.list_builds()
    ext.set_atoms(sy.q, sy.radius, sy.alpha)
    ext.set_covalent(sy.cov_atom, sy.cov_partner, sy.cov_strength)
    t2, f2, _ = ext.compute(sy.pos, sy.box)
    assert ext.list_builds() == builds + 1
    assert abs(E.total_energy(t2) / E.total_energy(t0) - 1) < 1e-12
    assert rel_rms(f2, f0) < 1e-11


def test_refused():
    sy = _water()
    i, j, t = _pairs(sy.pos, sy.box, sy.settings["cutoff"])
    eng = _engine(sy, "double")
    bad = t.copy()
    bad[0, 2] = 8
    with pytest.raises(E.PgmError, match="7 box vectors"):
        eng.set_candidates(*_segments(i, j, bad))
    jj = j.copy()
    jj[0] = sy.n
    with pytest.raises(E.PgmError, match="out of range"):
        eng.set_candidates(*_segments(i, jj, t))
    with pytest.raises(E.PgmError, match="itself"):
        eng.set_candidates([0], [[0, 0, 0]], [0, 1], [0])
    n = sy.n
    gv = np.zeros(n)
    eng.set_candidates(*_segments(i, j, t))
    with pytest.raises(E.PgmError, match="GVDW"):
        eng.set_gvdw(0, 0.5, gv + 1.0, gv, gv + 1.0)
    half = n // 3 * 3 // 2 // 3 * 3
    two = _engine(sy, "double")
    two.set_batch([0, half, n])
    with pytest.raises(E.PgmError, match="one system"):
        two.set_candidates(*_segments(i, j, t))
