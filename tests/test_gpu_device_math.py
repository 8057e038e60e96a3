"""The pair passes' own device math, evaluated on host arrays through the C ABI test hooks.

- FP32 erfc (`erfcFast` = `erfcxFast(x)` exp(-x^2)): exp(x^2) erfc(x) as t Q(t), t = 1/(1 + x/2), Q a degree-8
  polynomial fitted to it (pgm-md_amd/csrc/erfcx_fit.py), whose relative error is 1.6e-8 in exact arithmetic.
  In FP32 the exponent -x^2 carries rounding of order |x^2| 2^-24 (1.2e-6 at x = 4.5, the largest argument the
  kernels pass: the damping cut a r = 4.5; beta r stays below 3.2), and v_rcp_f32 / exp and the Horner steps
  add a few ulp: the bound checked against math.erfc is 3e-6 relative on [0, 4.5].
- FP64 pair kernels G0..G3 against the oracle's (the restatement of core/pgm_math.h, pinned to mpmath
  derivatives of erf(ar)/r in tests/test_oracle.py), with the damping cut the device kernels apply: beyond
  a r = 6 (FP64) or 4.5 (FP32) the erfc(a r) part is dropped, as the reference's GPU kernels do
  (openmm/platforms/common/src/CommonPGMKernels.cpp:351-352), so the reference here is the oracle's erfc form
  of beta alone there. Both sides run the same series / erfc branches: 1e-12 relative.
- FP32 pair kernels against the same reference with the FP32 cut, over the water pair types and the distances
  of the workloads (0.08 to 0.9 nm): FP32 arithmetic with a few-ulp erfc and up to three recurrence steps,
  bound 2e-5 relative per kernel.  The cut itself changes the kernels by at most 7.7e-6 relative (G3; 2.4e-7
  for G0) at these distances, checked against the uncut oracle.
- G0, G3 as the FP32 final and tangent passes rebuild them from G1, G2: G0 = r^2 G1 + h_a - h_b and
  G3 = (5 G2 + 4 b^4 h_b - 4 a^4 h_a)/r^2 (h_c = (2c/sqrt(pi)) exp(-c^2 r^2), h_a = 0 beyond the cut).  At
  large r the terms nearly cancel (r^2 G1 and h_b are ~20x G0 at b r = 3), so the FP32 error of G1, G2 is
  carried at the size of the terms: the bound is the same 2e-5, relative to the sum of the terms' magnitudes.
  A wrong coefficient or sign in either formula is an O(1) error against it.
"""

import math

import numpy as np
import pytest

import oracle as O
from pgmhip import engine as E

pytestmark = pytest.mark.gpu

BETA = 3.5
RADII = (0.075, 0.055)  # O, H of the water workloads (nm)


def _pairs():
    """(rg_sum, r) over the O-O, O-H and H-H types and 0.08..0.9 nm, denser where the branches switch."""
    rg = [2 * ri * ri + 2 * rj * rj for ri in RADII for rj in RADII]
    r = np.concatenate([np.linspace(0.08, 0.9, 1500), np.geomspace(0.08, 0.2, 500)])
    rs = np.repeat(rg, len(r))
    rr = np.tile(r, len(rg))
    return rs, rr


def test_erfc_fit_within_bound():
    x = np.linspace(0.0, 4.5, 200001, dtype=np.float32)
    y = E.eval_erfc_f32(x).astype(np.float64)
    ref = np.array([math.erfc(float(v)) for v in x])
    rel = np.abs(y / ref - 1.0)
    print(f"erfc fit: max relative error {rel.max():.3e} at x = {x[rel.argmax()]:.4f}")
    assert rel.max() < 3e-6
    # the fit alone (without FP32 rounding) is far tighter near 0, where the argument is small
    assert rel[x < 0.5].max() < 1e-6


def _reference(rs, rr, cut):
    """The oracle's G0..G3, with the erfc(a r) part dropped beyond a r = cut."""
    out = []
    for s, r in zip(rs, rr):
        a = 1.0 / math.sqrt(s)
        out.append(O.erfc_kernels(BETA, r) if a * r >= max(cut, 1.0) else O.pair_kernels(a, BETA, r))
    return np.array(out)


def test_damping_cut_effect_is_small():
    rs, rr = _pairs()
    full = np.array([O.pair_kernels(1.0 / math.sqrt(s), BETA, r) for s, r in zip(rs, rr)])
    for cut, bound in ((6.0, 1e-10), (4.5, 1e-5)):
        rel = np.abs(_reference(rs, rr, cut) - full) / np.abs(full)
        print(f"damping cut {cut}: max relative change per G_n {rel.max(0)}")
        assert rel.max() < bound


def test_fp64_pair_kernels_match_oracle():
    rs, rr = _pairs()
    G = E.eval_pair_kernels("double", rs, rr * rr, BETA)
    ref = _reference(rs, rr, 6.0)
    rel = np.abs(G[:, :4] - ref) / np.abs(ref)
    print("FP64 kernels vs oracle, max relative error per G_n:", rel.max(0))
    assert rel.max() < 1e-12
    assert np.array_equal(G[:, 4], G[:, 0]) and np.array_equal(G[:, 5], G[:, 3])


def test_fp32_pair_kernels_and_rebuilt_g0_g3():
    rs, rr = _pairs()
    r2 = rr * rr
    G32 = E.eval_pair_kernels("mixed", rs, r2, BETA)
    ref = _reference(rs, rr, 4.5)
    rel = np.abs(G32[:, :4] - ref) / np.abs(ref)
    print("FP32 kernels vs the oracle (FP32 cut), max relative error per G_n:", rel.max(0))
    assert rel.max() < 2e-5
    a = 1.0 / np.sqrt(rs)
    hb = 2.0 * BETA / math.sqrt(math.pi) * np.exp(-BETA * BETA * r2)
    ha = np.where(a * rr < 4.5, 2.0 * a / math.sqrt(math.pi) * np.exp(-a * a * r2), 0.0)
    terms0 = r2 * np.abs(ref[:, 1]) + hb + ha
    terms3 = (5.0 * np.abs(ref[:, 2]) + 4.0 * BETA**4 * hb + 4.0 * a**4 * ha) / r2
    rel0 = np.abs(G32[:, 4] - ref[:, 0]) / terms0
    rel3 = np.abs(G32[:, 5] - ref[:, 3]) / terms3
    print(f"FP32 rebuilt G0, G3 vs the oracle, relative to the terms: {rel0.max():.3e}, {rel3.max():.3e}; "
          f"relative to the value: {(np.abs(G32[:, 4] - ref[:, 0]) / np.abs(ref[:, 0])).max():.3e}, "
          f"{(np.abs(G32[:, 5] - ref[:, 3]) / np.abs(ref[:, 3])).max():.3e}")
    assert rel0.max() < 2e-5 and rel3.max() < 2e-5
