"""The CG preconditioners of the HIP engine (pgm_set_preconditioner) against the oracle's (tests/test_precond_oracle.py
pins those: M = D + D T_s D from the dense field tensor, the cap, the self term, convergence to the dense solve).

- FP64, 512 waters at tau 1e-10: the short-range preconditioner (radius 0.3 nm) takes the oracle's iterations
  with the same preconditioner, fewer than Jacobi, and ends at the oracle's residual; energies and forces are the
  oracle's.  Jacobi with the self term likewise.
- A radius that puts more than 48 pairs at an atom turns it off: Jacobi's iterations and results.
- Mixed precision: forces within the 5e-5 relative-RMS bound of the oracle.
- The adjoint's solve uses it too: the gradients equal the Jacobi engine's to the solver tolerance; a batch of
  uneven snapshots equals its single evaluations.
"""

import dataclasses

import numpy as np
import pytest

import oracle as O
from conftest import rel_rms
from pgmhip import engine as E
from pgmhip import systems as S

pytestmark = pytest.mark.gpu


def _water(tol=1e-10):
    sy = S.config0()
    sy.settings.update(tol=tol, max_iter=500, peek=0.0, predictor=0)
    return sy


def _oracle(sy, radius=0.0, jacobi_self=False):
    o = O.Oracle(sy)
    if radius or jacobi_self:
        o.set_preconditioner(radius, jacobi_self)
    e, f = o.compute(sy.pos, sy.box)
    return o, e, f


@pytest.mark.parametrize("radius, jacobi_self", [(0.3, False), (0.0, True)])
def test_fp64_matches_the_oracle(radius, jacobi_self):
    sy = _water()
    o, eo, fo = _oracle(sy, radius, jacobi_self)
    oj, _, _ = _oracle(sy)
    eng = E.Engine.for_system(sy, precision="double")
    eng.set_preconditioner(radius, jacobi_self)
    t, f, _ = eng.compute(sy.pos, sy.box)
    print(f"radius {radius} self {jacobi_self}: engine {eng.iterations()} iterations, oracle {o.iterations()}, "
          f"Jacobi {oj.iterations()}; residual {eng.residual():.3e} / {o.residual():.3e}")  # fmt: skip
    assert eng.iterations() == o.iterations()
    assert abs(eng.residual() / o.residual() - 1) < 1e-4
    if radius > 0:
        assert o.preconditioner_on() and o.iterations() < oj.iterations()
    assert abs(E.total_energy(t) / eo - 1) < 1e-11
    assert rel_rms(f, fo) < 1e-9


def test_cap_turns_it_off():
    sy = _water()
    jac = E.Engine.for_system(sy, precision="double")
    tj, fj, _ = jac.compute(sy.pos, sy.box)
    eng = E.Engine.for_system(sy, precision="double")
    eng.set_preconditioner(sy.settings["cutoff"])  # ~300 pairs per atom within the cutoff
    t, f, _ = eng.compute(sy.pos, sy.box)
    assert eng.iterations() == jac.iterations()
    assert abs(E.total_energy(t) / E.total_energy(tj) - 1) < 1e-12
    assert rel_rms(f, fj) < 1e-11


def test_mixed_within_the_force_bound():
    sy = _water(tol=1e-6)
    _, eo, fo = _oracle(_water())
    eng = E.Engine.for_system(sy, precision="mixed")
    eng.set_preconditioner(0.3)
    t, f, _ = eng.compute(sy.pos, sy.box)
    jac = E.Engine.for_system(sy, precision="mixed")
    jac.compute(sy.pos, sy.box)
    print(f"mixed tau 1e-6: {eng.iterations()} iterations (Jacobi {jac.iterations()}), forces {rel_rms(f, fo):.2e}")
    assert eng.iterations() <= jac.iterations()
    assert rel_rms(f, fo) < 5e-5
    assert abs(E.total_energy(t) / eo - 1) < 1e-6


def test_adjoint_uses_it():
    sy = _water()
    g = np.random.default_rng(4).normal(0.0, 1e-3, (sy.n, 3))
    out = []
    for radius in (0.0, 0.3):
        eng = E.Engine.for_system(sy, precision="double")
        eng.set_preconditioner(radius)
        eng.compute(sy.pos, sy.box)
        out.append((eng.backward(sy.pos, sy.box, g, 0.5), eng.adjoint_iterations()))
    (gj, itj), (gp, itp) = out
    print(f"adjoint iterations: Jacobi {itj}, short-range {itp}")
    assert itp < itj
    for k in gj:
        assert np.abs(gp[k] - gj[k]).max() <= 1e-8 * np.abs(gj[k]).max(), k


def test_batch_matches_single_systems():
    systems = [S.config3_snapshot(11), S.config3_snapshot(23), S.config3_snapshot(37)]
    for sy in systems:
        sy.settings.update(tol=1e-12, max_iter=2000, predictor=0, peek=0.0)
    from test_gpu_batch import _batch

    eng, off, pos, boxes = _batch(systems)
    eng.set_preconditioner(0.3)
    terms, f = eng.compute_batch(pos, boxes)
    for k, sy in enumerate(systems):
        single = E.Engine.for_system(dataclasses.replace(sy), precision="double")
        ts, fs, _ = single.compute(sy.pos, sy.box)
        assert abs(E.total_energy(terms[k]) / E.total_energy(ts) - 1) < 1e-11, k
        assert np.abs(f[off[k] : off[k + 1]] - fs).max() < 1e-8 * np.abs(fs).max(), k


def test_radius_checked():
    sy = _water()
    eng = E.Engine.for_system(sy, precision="double")
    with pytest.raises(E.PgmError, match="radius"):
        eng.set_preconditioner(sy.settings["cutoff"] + 0.1)
    with pytest.raises(E.PgmError, match="radius"):
        eng.set_preconditioner(-1.0)
