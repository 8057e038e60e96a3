"""Pins of the GVDW part of the oracle (oracle/pgm_oracle.cpp: gvdwG, gvdwPair, orc_gvdw_sum), the checker of the
engine's GVDW pair term.

* The definition: G(y) = F(y)/y^6, F = A^2 + g^2/2, A = erf y - (2y/sqrt(pi))(1 + 2y^2/3) e^{-y^2},
  g = (4y^3/(3 sqrt(pi))) e^{-y^2} (core/gvdw_kernel.h:4-20), evaluated with mpmath at 40 digits, and its
  derivative H = G'/y by mpmath differentiation; the 15 series coefficients used below y = 0.6 derived here with
  exact rational arithmetic from the series of erf and exp.
* The reference's own checks (core/tests/test_gvdw.cpp:25-50): G, H continuous across the switch at y = 0.6;
  dU/dr against central differences for both repulsion forms.
* The reference's pair function itself (core/gvdw_kernel.h compiled into oracle/_ref): the same numbers.
* The golden box (tests/golden/gvdw_box.npz from the reference's tests/data/gvdw_box.txt): pmemd-pgm's VDWAALS
  of 1536 atoms at rc = 10 A for the Slater and Gaussian forms, to 1e-3 kcal/mol (core/tests/test_gvdw.cpp:62-88).
* Forces and strain of the brute-force sum against central differences of its energy.
"""

import math
import os
from fractions import Fraction as Fr

import mpmath as mp
import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
HAVE_REF = os.path.exists(O.REF_LIB)
KCAL = 4.184


def _g_mp(y):
    y = mp.mpf(y)
    e = mp.e ** (-(y**2))
    A = mp.erf(y) - 2 / mp.sqrt(mp.pi) * y * (1 + 2 * y**2 / 3) * e
    g = 4 / (3 * mp.sqrt(mp.pi)) * y**3 * e
    return (A * A + g * g / 2) / y**6


def test_series_coefficients_from_the_definition():
    """G = C0 sum c_k y^{2k}, C0 = 8/(9 pi): with A = (2/sqrt(pi)) y alpha(y^2), G/C0 = (9/2) alpha^2/s^2 + e^{-2s}."""
    N = 20
    e = [Fr((-1) ** m, math.factorial(m)) for m in range(N)]
    erf = [Fr((-1) ** k, math.factorial(k) * (2 * k + 1)) for k in range(N)]
    alpha = [erf[k] - e[k] - (Fr(2, 3) * e[k - 1] if k else 0) for k in range(N)]
    a2 = [sum(alpha[i] * alpha[k - i] for i in range(k + 1)) for k in range(N)]
    assert a2[:4] == [0, 0, 0, 0]  # A = O(y^5): G is finite at 0
    c = [Fr(9, 2) * a2[k + 2] + Fr((-2) ** k, math.factorial(k)) for k in range(15)]
    assert c[0] == 1 and c[1] == -2 and c[2] == Fr(58, 25)
    c0 = 8 / (9 * math.pi)
    for y in (0.0, 0.1, 0.3, 0.45, 0.5999):
        G, H = O.gvdw_g(y)
        s = Fr(y) ** 2
        Gs = c0 * float(sum(ck * s**k for k, ck in enumerate(c)))
        Hs = c0 * float(sum(2 * k * ck * s ** (k - 1) for k, ck in enumerate(c) if k))
        assert abs(G - Gs) <= 1e-15 * abs(Gs), y
        assert abs(H - Hs) <= 1e-14 * abs(Hs), y


def test_g_and_h_against_mpmath():
    mp.mp.dps = 40
    for y in (0.05, 0.3, 0.59, 0.61, 0.8, 1.0, 1.5, 2.0, 3.0, 4.5, 6.0, 9.0):
        G, H = O.gvdw_g(y)
        Gm = _g_mp(y)
        Hm = mp.diff(_g_mp, y) / y
        assert abs(G - float(Gm)) <= 1e-13 * abs(float(Gm)), y
        assert abs(H - float(Hm)) <= 1e-11 * abs(float(Hm)), y
    assert O.gvdw_g(0.0)[0] == pytest.approx(8 / (9 * math.pi), rel=1e-15)


def test_reference_checks_switch_and_finite_differences():
    G1, H1 = O.gvdw_g(0.6 - 1e-12)
    G2, H2 = O.gvdw_g(0.6 + 1e-12)
    assert abs(G1 - G2) <= 1e-9 * abs(G2) and abs(H1 - H2) <= 1e-7 * abs(H2)
    beta = 1.0 / math.sqrt(2.0 * (0.07**2 + 0.07**2))
    for rep, B in ((0, 0.9453), (1, 4.52)):
        for r in (0.02, 0.05, 0.1, 0.2, 0.28, 0.35, 0.5, 0.9):
            _, d = O.gvdw_pair(r, beta, 800.0, 0.0025, B, rep)
            h = 1e-6
            fd = -(O.gvdw_pair(r + h, beta, 800.0, 0.0025, B, rep)[0] - O.gvdw_pair(r - h, beta, 800.0, 0.0025, B, rep)[0])
            fd /= 2 * h * r
            assert abs(fd - d) <= 1e-6 * max(abs(fd), 1e-12), (rep, r)


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (reference sources absent)")
def test_pair_function_equals_the_reference():
    rng = np.random.default_rng(5)
    for _ in range(400):
        r, bij = rng.uniform(0.0, 1.2), rng.uniform(3.0, 12.0)
        A, C6, B, rep = rng.uniform(0, 1e4), rng.uniform(0, 0.01), rng.uniform(0.5, 5.0), int(rng.integers(2))
        e, d = O.gvdw_pair(r, bij, A, C6, B, rep)
        er, dr = O.gvdw_pair(r, bij, A, C6, B, rep, kind="reference")
        assert abs(e - er) <= 1e-14 * max(abs(er), 1e-300) + 1e-300
        assert abs(d - dr) <= 1e-13 * max(abs(dr), 1e-300) + 1e-300


def golden_box():
    return np.load(os.path.join(HERE, "golden", "gvdw_box.npz"))


def box_parameters(g, k):
    """Per-atom sqrt(A), sqrt(C6), b in kJ/mol nm units for model k of the golden box (GVDW atoms: has_lj)."""
    lj = g["has_lj"].astype(float)
    A = g["model_arep_kcal"][k] * KCAL
    C6 = g["model_c6_kcal_a6"][k] * KCAL * 1e-6
    return np.sqrt(A) * lj, np.sqrt(C6) * lj, np.full(len(lj), g["model_b"][k])


@pytest.mark.parametrize("k", [0, 1])
def test_golden_box_matches_pmemd(k):
    g = golden_box()
    sa, sc, b = box_parameters(g, k)
    E, _, _ = O.gvdw_sum(g["pos"], g["box"], g["radius"], sa, sc, b, int(g["model_rep"][k]), float(g["rc"]),
                         g["excluded"][:, 0], g["excluded"][:, 1])  # fmt: skip
    print(f"model {k}: {E / KCAL:.6f} kcal/mol, pmemd-pgm {g['pmemd_kcal'][k]}")
    assert abs(E / KCAL - g["pmemd_kcal"][k]) < 1e-3


def test_sum_forces_and_strain_by_finite_differences():
    g = golden_box()
    sel = np.arange(0, 300)
    pos, box, R = g["pos"][sel], g["box"], g["radius"][sel]
    sa, sc, b = (v[sel] for v in box_parameters(g, 1))
    ex = g["excluded"]
    keep = (ex < 300).all(1)
    si, sj = ex[keep, 0], ex[keep, 1]
    rep, rc = 0, 0.9

    def energy(p, H=box):
        return O.gvdw_sum(p, H, R, sa, sc, b, rep, rc, si, sj)[0]

    E, F, S = O.gvdw_sum(pos, box, R, sa, sc, b, rep, rc, si, sj)
    h = 1e-6
    for i in (0, 7, 150, 299):
        for a in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, a] += h
            p2[i, a] -= h
            fd = -(energy(p1) - energy(p2)) / (2 * h)
            assert abs(fd - F[i, a]) <= 1e-5 * np.abs(F).max(), (i, a)
    # strain: E(x (1 + eps)) with the box deformed alike, S_ab = dE/d eps_ab
    for a, c in ((0, 0), (1, 1), (0, 1), (2, 0)):
        D = np.zeros((3, 3))
        D[a, c] = h
        e1 = energy(pos @ (np.eye(3) + D).T, box @ (np.eye(3) + D).T)
        e2 = energy(pos @ (np.eye(3) - D).T, box @ (np.eye(3) - D).T)
        assert abs((e1 - e2) / (2 * h) - S[a, c]) <= 1e-5 * np.abs(S).max(), (a, c)
