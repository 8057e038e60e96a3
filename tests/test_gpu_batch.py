"""Batches of independent systems in one engine (pgm_set_batch) and the pair-pass autotuner.

A batch must give every system the energy terms, forces and parameter gradients of that system evaluated
alone: the same kernels run over all systems in one launch, and the CG solves the block-diagonal system of all
of them, so with the solves converged to tau = 1e-12 the batch and the single evaluations agree to FP64
round-off and solver tolerance.  Each system is also checked against the FP64 oracle.
"""

import dataclasses

import numpy as np
import pytest

import oracle as O
from pgmhip import engine as E
from pgmhip import systems as S

pytestmark = pytest.mark.gpu

TIGHT = dict(tol=1e-12, max_iter=2000, predictor=0, peek=0.0)


def _triclinic_snapshot(nmol=220, seed=4):
    """Random rigid waters in a reduced-triclinic box wide enough for the 0.9 nm cutoff."""
    box = np.array([[2.7, 0.0, 0.0], [0.5, 2.6, 0.0], [-0.4, 0.6, 2.5]])
    rng = np.random.default_rng(seed)
    tmpl = S.water_template()
    R = S.random_rotations(rng, nmol)
    centers = rng.random((nmol, 3)) @ box
    pos = (centers[:, None, :] + np.einsum("mij,aj->mai", R, tmpl)).reshape(-1, 3)
    base = S.water_box(nmol, 2.5, 8, seed, 0.0, "tri", S.base_settings(grid=(30, 30, 30), tol=1e-8))
    return dataclasses.replace(base, name=f"tri{nmol}", pos=pos, box=box)


def _systems():
    sys_ = [S.config3_snapshot(41), S.config3_snapshot(57), _triclinic_snapshot(), S.config3_snapshot(18)]
    for s in sys_:
        s.settings.update(TIGHT)
    return sys_


def _batch(systems, precision="double"):
    n = [s.n for s in systems]
    off = np.concatenate([[0], np.cumsum(n)]).astype(np.int32)
    cat = lambda k: np.concatenate([getattr(s, k) for s in systems])  # noqa: E731
    ca = np.concatenate([s.cov_atom + off[i] for i, s in enumerate(systems)])
    cp = np.concatenate([s.cov_partner + off[i] for i, s in enumerate(systems)])
    st = dict(systems[0].settings)
    keys = ("cutoff", "beta", "order", "tol", "max_iter", "peek", "predictor", "ke", "skin", "polarizable")
    eng = E.Engine(precision=precision, grid=tuple(systems[0].settings["grid"]), **{k: st[k] for k in keys})
    eng.set_atoms(cat("q"), cat("radius"), cat("alpha"))
    eng.set_covalent(ca, cp, cat("cov_strength"))
    eng.set_batch(off, [s.settings["grid"] for s in systems])
    boxes = np.stack([s.box for s in systems])
    return eng, off, np.concatenate([s.pos for s in systems]), boxes


def test_batch_matches_single_systems_and_oracle():
    systems = _systems()
    eng, off, pos, boxes = _batch(systems)
    assert eng.num_systems() == len(systems)
    terms, f = eng.compute_batch(pos, boxes)
    for k, sy in enumerate(systems):
        fk = f[off[k] : off[k + 1]]
        single = E.Engine.for_system(sy, precision="double")
        ts, fs, _ = single.compute(sy.pos, sy.box)
        es, eb = E.total_energy(ts), E.total_energy(terms[k])
        assert abs(eb / es - 1) < 1e-11, (k, eb, es)
        assert np.abs(fk - fs).max() < 1e-8 * np.abs(fs).max(), k
        o = O.Oracle(sy)
        eo, fo = o.compute(sy.pos, sy.box)
        assert abs(eb / eo - 1) < 1e-11, (k, eb, eo)
        assert E.rel_rms(fk, fo) < 1e-9, k


def test_batch_mixed_within_force_bound():
    systems = _systems()
    eng64, off, pos, boxes = _batch(systems)
    t64, f64 = eng64.compute_batch(pos, boxes)
    engm, _, _, _ = _batch(systems, "mixed")
    tm, fm = engm.compute_batch(pos, boxes)
    for k in range(len(systems)):
        sl = slice(off[k], off[k + 1])
        assert E.rel_rms(fm[sl], f64[sl]) < 5e-5
        assert abs(E.total_energy(tm[k]) / E.total_energy(t64[k]) - 1) < 1e-6


def test_batch_backward_matches_single_systems():
    systems = _systems()
    eng, off, pos, boxes = _batch(systems)
    rng = np.random.default_rng(3)
    g = rng.normal(0.0, 1e-3, pos.shape)
    wE = np.array([0.1, -0.2, 0.3, 0.05])
    eng.compute_batch(pos, boxes)
    gb = eng.backward_batch(pos, boxes, g, wE)
    assert eng.adjoint_iterations() > 0
    coff = np.concatenate([[0], np.cumsum([len(s.cov_atom) for s in systems])])
    for k, sy in enumerate(systems):
        single = E.Engine.for_system(sy, precision="double")
        single.compute(sy.pos, sy.box)
        gs = single.backward(sy.pos, sy.box, g[off[k] : off[k + 1]], wE[k])
        for name in ("q", "radius", "alpha"):
            a, b = gb[name][off[k] : off[k + 1]], gs[name]
            assert np.abs(a - b).max() < 1e-8 * np.abs(b).max(), (k, name)
        a, b = gb["cov_strength"][coff[k] : coff[k + 1]], gs["cov_strength"]
        assert np.abs(a - b).max() < 1e-8 * np.abs(b).max(), (k, "cov")


def test_batch_errors_and_reset():
    systems = _systems()[:2]
    eng, off, pos, boxes = _batch(systems)
    with pytest.raises(E.PgmError, match="batch"):
        eng.compute(pos, boxes[0])
    with pytest.raises(E.PgmError, match="offsets"):
        eng.set_batch([0, 10, off[-1] + 1])
    with pytest.raises(E.PgmError, match="at least one atom"):
        eng.set_batch([0, 10, 10, off[-1]])
    with pytest.raises(E.PgmError, match="two systems"):
        eng.set_covalent([0], [off[1]], [0.01])
    # another atom count: one system again
    sy = systems[0]
    eng.set_atoms(sy.q, sy.radius, sy.alpha)
    eng.set_covalent(sy.cov_atom, sy.cov_partner, sy.cov_strength)
    assert eng.num_systems() == 1
    eng.compute(sy.pos, sy.box)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_autotune_keeps_results(precision):
    """Every pair-pass variant (fused or separate compaction, 16/32/64 lanes, in FP32 the field pass on the fly or
    from stored kernels) gives the same results (FP64: to round-off; mixed: to FP32 summation order); the
    autotuner reports its timings and picks the fastest."""
    sy = S.config3_snapshot(41)
    sy.settings.update(TIGHT)
    eng = E.Engine.for_system(sy, precision=precision)
    t0, f0, _ = eng.compute(sy.pos, sy.box)
    tol_e, tol_f = (1e-13, 1e-10) if precision == "double" else (1e-7, 1e-5)
    for lanes in (16, 32, 64):
        for fused in (True, False):
            for otf in ((False, True) if precision == "mixed" else (False,)):
                eng.set_lanes(lanes, fused)
                if precision == "mixed":
                    eng.set_field_otf(otf)
                t, f, _ = eng.compute(sy.pos, sy.box, new_step=False)
                assert abs(E.total_energy(t) / E.total_energy(t0) - 1) < tol_e, (lanes, fused, otf)
                assert np.abs(f - f0).max() < tol_f * np.abs(f0).max(), (lanes, fused, otf)
    best, ms = eng.autotune(2)
    print(f"autotune {precision} (ms per evaluation in the pair passes):", ms, "->", best, "setup", eng.setup_lanes())
    assert best[0] in (16, 32, 64) and all(v > 0 for v in ms.values())
    assert ms[best] == min(ms.values())
    assert eng.setup_lanes() in (16, 32, 64)
    t, f, _ = eng.compute(sy.pos, sy.box, new_step=False)
    assert np.abs(f - f0).max() < tol_f * np.abs(f0).max()
    # the setup pass at a lane count of its own, and set_lanes setting both again
    other = 64 if best[0] != 64 else 16
    eng.set_setup_lanes(other)
    assert eng.setup_lanes() == other
    t, f, _ = eng.compute(sy.pos, sy.box, new_step=False)
    assert abs(E.total_energy(t) / E.total_energy(t0) - 1) < tol_e
    assert np.abs(f - f0).max() < tol_f * np.abs(f0).max()
    eng.set_lanes(best[0], best[1])
    assert eng.setup_lanes() == best[0]
    with pytest.raises(E.PgmError):
        eng.set_setup_lanes(48)
