"""A pGM dump (GROMACS GMX_PGM_DUMP layout) recomputed by the HIP engine: the dump of an oracle evaluation of
512 waters (tau 1e-10, no predictor) written and read back, then evaluated by the FP64 engine through
pgmhip.dump.recompute, as core/tests/run_dump.cpp does with the reference core. FP64 bounds: the engine matches
the oracle to round-off and the solver tolerance."""

import numpy as np
import pytest

import oracle as O
from pgmhip import dump as D
from pgmhip import systems as S

pytestmark = pytest.mark.gpu


def test_engine_recomputes_a_dump(tmp_path):
    sy = S.config0()
    sy.settings.update(tol=1e-10, max_iter=500, peek=0.0, predictor=0)
    o = O.Oracle(sy)
    _, f = o.compute(sy.pos, sy.box)
    d = D.from_evaluation(sy, o.terms(), f, o.strain(), o.induced(), o.permanent(), o.iterations(), o.residual())
    p = tmp_path / "w512.dump"
    D.write(str(p), d)
    r = D.recompute(D.read(str(p)))
    print(r)
    for k, (a, b) in r["terms"].items():
        assert abs(a - b) <= 1e-11 * abs(a) + 1e-9, k
    assert abs(r["total_rel"]) < 1e-12
    assert r["force_max_diff"] < 1e-9 * r["force_rms"]
    assert r["induced_max_diff"] < 1e-10 and r["permanent_max_diff"] < 1e-14
    assert r["virial_max_diff"] < 1e-9 * r["virial_max"]
    assert abs(r["iterations"][0] - r["iterations"][1]) <= 1
