"""pgm_field, the pgmcore::Engine::field operator (core/pgm_core.h:76-79, core/pgm_core.cpp:268-296): the field of
given charges and dipoles at every atom (direct + reciprocal + self), against the oracle's field at the same
geometry, and its linearity in (q, d)."""

import numpy as np
import pytest

import oracle as O
from conftest import rel_rms
from pgmhip import engine as E
from pgmhip import systems as S

pytestmark = pytest.mark.gpu


def _operands(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, 0.4, n), rng.normal(0.0, 0.01, (n, 3))


@pytest.mark.parametrize("maker", [S.config0, S.config1], ids=["water512", "ring8550"])
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_field_matches_the_oracle(maker, precision):
    sy = maker()
    q, d = _operands(sy.n, 3)
    eng = E.Engine.for_system(sy, precision=precision)
    Eg = eng.field(sy.pos, sy.box, q, d)
    o = O.Oracle(sy)
    o.set_geometry(sy.pos, sy.box)
    Eo = o.field(q, d)
    if precision == "double":
        assert np.abs(Eg - Eo).max() <= 1e-11 * np.abs(Eo).max()
    else:
        assert rel_rms(Eg, Eo) < 2e-6


def test_field_is_linear_and_leaves_evaluations_alone():
    sy = S.config0()
    eng = E.Engine.for_system(sy, precision="double")
    t0, f0, _ = eng.compute(sy.pos, sy.box)
    q1, d1 = _operands(sy.n, 4)
    q2, d2 = _operands(sy.n, 5)
    E1 = eng.field(sy.pos, sy.box, q1, d1)
    E2 = eng.field(sy.pos, sy.box, q2, d2)
    E12 = eng.field(sy.pos, sy.box, q1 + 2.0 * q2, d1 + 2.0 * d2)
    assert np.abs(E12 - (E1 + 2.0 * E2)).max() <= 1e-11 * np.abs(E12).max()
    # charges only: the charge field; the engine's own evaluation afterwards is unchanged
    Eq = eng.field(sy.pos, sy.box, sy.q, np.zeros((sy.n, 3)))
    assert np.isfinite(Eq).all() and np.abs(Eq).max() > 0
    t1, f1, _ = eng.compute(sy.pos, sy.box, new_step=False)
    assert abs(E.total_energy(t1) - E.total_energy(t0)) <= 1e-9 * abs(E.total_energy(t0))
