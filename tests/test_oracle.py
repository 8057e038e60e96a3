"""Pins of the FP64 CPU oracle (oracle/pgm_oracle.cpp) before it is trusted as the checker of the HIP engine:

* the reference's golden vectors: pGM-JAX FP64 energies, forces, induced dipoles and iteration counts of
  512 waters at the four levels q / qp / qi / qpi (reference tests/data/ref_w512_*_novdw.npz, published
  agreement docs/openmm.md:102-111);
* the reference itself, compiled from core/pgm_core.cpp by oracle/Makefile (oracle/_ref), on the same
  inputs: whole evaluations, fields, strain, predictor trajectories and every numerics routine;
* mathematics that fixes the answer independently of both: finite differences of the energy (forces and
  strain, the core/tests/test_core.cpp pattern), the explicit induction matrix solved densely, an exact
  Ewald sum for the reciprocal energy, high-precision radial kernels, B-spline moduli from the explicit
  spline, numpy's FFT.
"""

import os

import numpy as np
import pytest

import oracle as O
from conftest import golden_system, rel_rms, triclinic_waters
from pgmhip import systems as S

HAVE_REF = os.path.exists(O.REF_LIB)
needs_ref = pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (reference sources absent)")


# ---------------------------------------------------------------------------------------------- golden


@pytest.mark.parametrize("level", ["q", "qp", "qi", "qpi"])
def test_golden_pgmjax(golden, level):
    sy = golden_system(golden, level)
    o = O.Oracle(sy)
    e, f = o.compute(sy.pos, sy.box)
    er = float(golden[f"{level}_energy"])
    assert abs(e / er - 1) < 1e-12
    assert np.abs(f - golden[f"{level}_forces"]).max() < 2e-9  # kJ/mol/nm, |F| ~ 1e3
    assert np.abs(o.induced() - golden[f"{level}_mu"]).max() < 1e-13
    assert o.iterations() == int(golden[f"{level}_iterations"])


# ---------------------------------------------------------------------------------------------- reference


def _compare_engines(sy, pos_list, box, kind_a="port", kind_b="reference", keys=None, tol_e=1e-12, tol_f=1e-9):
    a, b = O.Oracle(sy, kind_a), O.Oracle(sy, kind_b)
    for k, pos in enumerate(pos_list):
        if keys is not None:
            a.set_history_key(keys[k])
            b.set_history_key(keys[k])
        ea, fa = a.compute(pos, box)
        eb, fb = b.compute(pos, box)
        assert abs(ea - eb) <= tol_e * abs(eb)
        assert np.abs(fa - fb).max() <= tol_f * np.abs(fb).max()
        assert a.iterations() == b.iterations()
        ta, tb = a.terms(), b.terms()
        for t in ta:
            assert abs(ta[t] - tb[t]) <= tol_e * max(1.0, abs(eb))
        np.testing.assert_allclose(a.strain(), b.strain(), rtol=0, atol=tol_f * np.abs(b.strain()).max())
        np.testing.assert_allclose(a.induced(), b.induced(), rtol=0, atol=1e-12)
        np.testing.assert_allclose(a.permanent(), b.permanent(), rtol=0, atol=1e-15)
    return a, b


@needs_ref
def test_reference_config0():
    sy = S.config0()
    _compare_engines(sy, [sy.pos], sy.box)


@needs_ref
def test_reference_triclinic_and_nocutoff():
    sy = triclinic_waters(24, 7)
    _compare_engines(sy, [sy.pos], sy.box)
    sy0 = triclinic_waters(10, 7, periodic=False)
    _compare_engines(sy0, [sy0.pos], None)


@needs_ref
def test_reference_predictor_trajectory():
    """mu4 predictor, peek and the same-step rule over a short trajectory (core/pgm_core.cpp:307-328,
    386-397): identical iteration counts and dipoles."""
    sy = S.config0()
    sy.settings.update(tol=1e-5, predictor=2, peek=0.65)
    traj = S.rigid_moves(sy, 3, 0.002, 6)
    pos = [traj[0], traj[1], traj[1], traj[2], traj[3], traj[4], traj[5], traj[5]]
    keys = [0, 1, 1, 2, 3, 4, 5, 5]
    a, b = _compare_engines(sy, pos, sy.box, keys=keys, tol_e=1e-11, tol_f=1e-8)


@needs_ref
def test_reference_field():
    sy = triclinic_waters(12, 3)
    rng = np.random.default_rng(0)
    q = rng.normal(0, 0.5, sy.n)
    d = rng.normal(0, 0.02, (sy.n, 3))
    a, b = O.Oracle(sy, "port"), O.Oracle(sy, "reference")
    a.set_geometry(sy.pos, sy.box)
    b.set_geometry(sy.pos, sy.box)
    np.testing.assert_allclose(a.field(q, d), b.field(q, d), rtol=0, atol=1e-11)


@needs_ref
def test_reference_numerics():
    for a, beta, r in [(8.7, 3.5, 0.05), (8.7, 3.5, 0.1149), (8.7, 3.5, 0.115), (6.67, 3.5, 0.5), (9.1, 3.5, 0.89),
                       (3.0, 5.0, 0.3), (12.0, 3.0, 1e-4)]:  # fmt: skip
        for fn in ("pair_kernels", "erf_kernels"):
            args = (a, beta, r) if fn == "pair_kernels" else (a, r)
            np.testing.assert_array_equal(getattr(O, fn)(*args, kind="port"), getattr(O, fn)(*args, kind="reference"))
        np.testing.assert_array_equal(O.erfc_kernels(a, r, kind="port"), O.erfc_kernels(a, r, kind="reference"))
    for K, p in [(32, 6), (36, 6), (48, 6), (72, 6), (20, 5), (24, 8)]:
        np.testing.assert_allclose(O.bspline_moduli(K, p), O.bspline_moduli(K, p, kind="reference"), rtol=1e-13)
    assert O.ewald_coefficient(0.9, 5e-4) == O.ewald_coefficient(0.9, 5e-4, kind="reference")
    rng = np.random.default_rng(1)
    for shape in [(6, 10, 8), (12, 9, 15), (36, 36, 36)]:
        x = rng.normal(size=shape)
        F = O.fft3d_r2c(x)
        np.testing.assert_allclose(F, O.fft3d_r2c(x, kind="reference"), rtol=0, atol=1e-10)
        np.testing.assert_allclose(O.fft3d_c2r(F, shape[2]), O.fft3d_c2r(F, shape[2], kind="reference"), atol=1e-9)


# ---------------------------------------------------------------------------------------------- mathematics


def test_fft_vs_numpy():
    rng = np.random.default_rng(2)
    for shape in [(6, 10, 8), (12, 9, 15), (20, 14, 7), (32, 32, 32)]:
        x = rng.normal(size=shape)
        F = O.fft3d_r2c(x)
        np.testing.assert_allclose(F, np.fft.rfftn(x), rtol=0, atol=1e-10)
        np.testing.assert_allclose(O.fft3d_c2r(F, shape[2]), np.fft.irfftn(F, s=shape, axes=(0, 1, 2)) * x.size, atol=1e-9)


def test_radial_kernels_high_precision():
    """B_{n+1} = -(1/r) dB_n/dr of erf(a r)/r by mpmath differentiation at 40 digits; G_n = B_n(a) - B_n(beta)
    on both sides of the series/closed-form switch at a r = 1 and far into the erfc branch."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40

    def B(a, n):
        f = [lambda r: mp.erf(a * r) / r]
        for k in range(n):
            f.append((lambda g: (lambda r: -mp.diff(g, r) / r))(f[-1]))
        return f

    for a, beta, r in [(8.7, 3.5, 0.05), (8.7, 3.5, 0.1149), (8.7, 3.5, 0.1151), (6.67, 3.5, 0.4), (7.6, 3.5, 0.85)]:
        Ba, Bb = B(mp.mpf(a), 3), B(mp.mpf(beta), 3)
        G = O.pair_kernels(a, beta, r)
        for n in range(4):
            exact = float(Ba[n](mp.mpf(r)) - Bb[n](mp.mpf(r)))
            assert abs(G[n] - exact) <= 1e-12 * abs(exact) + 1e-300, (a, r, n)
    # finite at r -> 0: G_n(0) = (2/sqrt(pi)) (2^n/(2n+1)) (a^(2n+1) - beta^(2n+1))
    a, beta = 8.7, 3.5
    G = O.pair_kernels(a, beta, 1e-9)
    for n in range(4):
        lim = 2 / np.sqrt(np.pi) * 2**n / (2 * n + 1) * (a ** (2 * n + 1) - beta ** (2 * n + 1))
        assert abs(G[n] / lim - 1) < 1e-10


def test_bspline_moduli_explicit():
    from math import comb, factorial

    def M(p, x):
        return sum((-1) ** k * comb(p, k) * max(x - k, 0.0) ** (p - 1) for k in range(p + 1)) / factorial(p - 1)

    for K, p in [(32, 6), (36, 6), (20, 5), (24, 8), (72, 6)]:
        c = np.array([M(p, k + 1) for k in range(K)])
        b = np.abs(np.fft.ifft(c) * K) ** 2
        got = O.bspline_moduli(K, p)
        ok = b > 1e-7
        np.testing.assert_allclose(got[ok], b[ok], rtol=1e-12)
        for m in np.where(~ok)[0]:
            assert got[m] == pytest.approx(0.5 * (got[(m - 1) % K] + got[(m + 1) % K]))


def test_ewald_coefficient():
    from math import erfc

    for rc, tol in [(0.9, 5e-4), (1.2, 1e-6), (0.8, 1e-5)]:
        beta = O.ewald_coefficient(rc, tol)
        assert abs(erfc(beta * rc) / tol - 1) < 1e-12


def _ewald_reciprocal(pos, box, q, d, beta, nmax):
    """Exact Ewald reciprocal energy [e^2/nm] of point charges and dipoles:
    (1/(2 pi V)) sum_{m != 0} exp(-pi^2 m^2/beta^2)/m^2 |sum_j (q_j + 2 pi i m.d_j) exp(2 pi i m.r_j)|^2."""
    recip = np.linalg.inv(box)  # columns: reciprocal vectors
    V = abs(np.linalg.det(box))
    n = np.arange(-nmax, nmax + 1)
    N = np.stack(np.meshgrid(n, n, n, indexing="ij"), -1).reshape(-1, 3)
    N = N[np.any(N != 0, axis=1)]
    m = N @ recip.T
    m2 = (m * m).sum(1)
    ph = np.exp(2j * np.pi * (m @ pos.T))
    Sm = (ph * (q[None, :] + 2j * np.pi * (m @ d.T))).sum(1)
    return float((np.exp(-np.pi**2 * m2 / beta**2) / m2 * np.abs(Sm) ** 2).sum() / (2 * np.pi * V))


@pytest.mark.parametrize("triclinic", [False, True])
def test_pme_reciprocal_vs_exact_ewald(triclinic):
    """The tiny-box cases of the reference's tests/pme_tiny.py (charges, a dipole, both; cubic and triclinic
    2 nm boxes) against the exact Ewald sum: order 8, 40^3 grid."""
    from pgmhip.systems import System

    box = np.array([[2.0, 0, 0], [-0.6, 1.9, 0], [-0.6, -0.9, 1.7]]) if triclinic else np.diag([2.0, 2.0, 2.0])
    pos = np.array([[0.31, 0.42, 0.53], [1.11, 0.97, 1.23], [0.5, 1.4, 0.2]])
    for q, cov in (([1.0, -1.0, 0.0], []), ([0.0, 0.0, 0.0], [(2, 0, 0.05)]),
                   ([0.7, -0.4, -0.3], [(2, 0, 0.05), (0, 1, -0.03)])):  # fmt: skip
        cov = np.array(cov).reshape(-1, 3)
        st = dict(periodic=1, cutoff=0.05, beta=3.5, grid=(40, 40, 40), order=8, polarizable=0, tol=1e-8, peek=0.0,
                  predictor=0, ke=1.0)  # fmt: skip
        sy = System("tiny", pos, box, np.array(q), np.full(3, 0.07), np.zeros(3), cov[:, 0].astype(np.int32),
                    cov[:, 1].astype(np.int32), cov[:, 2].astype(float), st)  # fmt: skip
        o = O.Oracle(sy)
        o.compute(pos, box)
        d = o.permanent()
        exact = _ewald_reciprocal(pos, box, np.array(q), d, 3.5, 12)
        # order-8 PME discretization error at 0.05 nm spacing is ~3e-9 relative
        assert abs(o.terms()["reciprocal"] - exact) < 2e-8 * max(1.0, abs(exact))


def test_total_energy_independent_of_beta():
    """The Ewald split: with erfc(beta rc) negligible and a fine grid, the total does not depend on beta."""
    sy = triclinic_waters(16, 11)
    energies = []
    for beta in (4.5, 5.0, 5.5):
        o = O.Oracle(sy, beta=beta, grid=(64, 64, 64), order=8, cutoff=0.85, tol=1e-12)
        energies.append(o.compute(sy.pos, sy.box, forces=False)[0])
    assert np.ptp(energies) < 2e-6 * abs(energies[0])


def test_induction_dense_solve():
    """(alpha^-1 - T) mu = E_perm solved densely with T assembled column by column from the field of unit
    dipoles; T symmetric, the matrix positive definite, the CG solution equal to the dense one."""
    sy = triclinic_waters(8, 5)
    o = O.Oracle(sy, tol=1e-14, max_iter=5000)
    o.compute(sy.pos, sy.box, forces=False)
    mu = o.induced()
    n = sy.n
    zq = np.zeros(n)
    T = np.zeros((3 * n, 3 * n))
    for k in range(3 * n):
        e = np.zeros(3 * n)
        e[k] = 1.0
        T[:, k] = o.field(zq, e.reshape(n, 3)).reshape(-1)
    np.testing.assert_allclose(T, T.T, atol=1e-9 * np.abs(T).max())
    A = np.diag(np.repeat(1.0 / sy.alpha, 3)) - T
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0
    b = o.field(sy.q, o.permanent()).reshape(-1)
    np.testing.assert_allclose(mu.reshape(-1), np.linalg.solve(A, b), rtol=0, atol=1e-10 * np.abs(mu).max())


@pytest.mark.parametrize("periodic", [True, False])
def test_forces_and_strain_finite_differences(periodic):
    """core/tests/test_core.cpp:40-110: central differences, h = 1e-5 nm (forces) and 1e-6 (strain)."""
    sy = triclinic_waters(24 if periodic else 10, 7, periodic=periodic)
    box = sy.box if periodic else None
    o = O.Oracle(sy)
    e0, F = o.compute(sy.pos, box)
    h, err, fmax = 1e-5, 0.0, 0.0
    for i in (0, 1, 2, 7, 20):
        for k in range(3):
            p = sy.pos.copy()
            p[i, k] += h
            ep = O.Oracle(sy).compute(p, box, forces=False)[0]
            p[i, k] -= 2 * h
            em = O.Oracle(sy).compute(p, box, forces=False)[0]
            err = max(err, abs(-(ep - em) / (2 * h) - F[i, k]))
            fmax = max(fmax, abs(F[i, k]))
    assert err <= 1e-5 * fmax
    if not periodic:
        return
    S_ = o.strain()
    hs, errS = 1e-6, 0.0
    for a in range(3):
        for b in range(a, 3):
            ef = []
            for eps in (hs, -hs):
                p = sy.pos.copy()
                p[:, a] += eps * p[:, b]
                bx = sy.box.copy()
                bx[:, a] += eps * bx[:, b]
                ef.append(O.Oracle(sy).compute(p, bx, forces=False)[0])
            errS = max(errS, abs((ef[0] - ef[1]) / (2 * hs) - S_[a, b]))
    assert errS <= 1e-5 * np.abs(S_).max()


def test_same_key_rule():
    """One history entry per key: a second evaluation with the same key starts from that solution and leaves
    the history alone (core/pgm_core.h:56-61)."""
    sy = S.config0()
    o = O.Oracle(sy, tol=1e-6, predictor=2, peek=0.0)
    o.set_history_key(1)
    o.compute(sy.pos, sy.box, forces=False)
    assert o.iterations() > 0
    o.compute(sy.pos, sy.box, forces=False)
    assert o.iterations() == 0  # same key: starts from the converged solution


def test_rel_rms_definition():
    f = np.array([[1.0, 2.0, 3.0]])
    assert rel_rms(f * 1.01, f) == pytest.approx(0.01)
