"""The seeded workload generators and the prmtop reader (host-side logic)."""

import os

import numpy as np
import pytest

from pgmhip import prmtop
from pgmhip import systems as S


def _min_image(d, L):
    return d - L * np.round(d / L)


def test_water_boxes():
    for maker, n, L in [(S.config0, 1536, 2.48), (S.config2, 96000, 6.84)]:
        sy = maker()
        assert sy.n == n and np.allclose(sy.box, np.diag([L] * 3))
        o, h1, h2 = sy.pos[0::3], sy.pos[1::3], sy.pos[2::3]
        assert np.allclose(np.linalg.norm(h1 - o, axis=1), 0.09572)
        cosang = ((h1 - o) * (h2 - o)).sum(1) / 0.09572**2
        assert np.allclose(np.degrees(np.arccos(cosang)), 104.52)
        assert np.isclose(sy.q.sum(), 0.0) and len(sy.cov_atom) == 2 * n // 3
        assert np.all(sy.cov_atom % 3 == 0) and np.all(sy.cov_strength == 0.01)
        np.testing.assert_array_equal(maker().pos, sy.pos)  # seeded
    assert S.config2().n / 6.84**3 == pytest.approx(300.0, rel=0.01)


def test_ring_liquid():
    sy = S.config1()
    assert sy.n == 8550 and sy.box[0, 0] == 4.5
    q = sy.q.reshape(450, 19)
    assert np.allclose(q.sum(1), 0.0)
    assert np.all((sy.alpha >= 0.3e-3) & (sy.alpha <= 1.8e-3))
    assert np.all((sy.radius >= 0.05) & (sy.radius <= 0.09))
    assert np.all((sy.cov_strength >= 0) & (sy.cov_strength <= 0.015))
    d = np.linalg.norm(_min_image(sy.pos[sy.cov_partner] - sy.pos[sy.cov_atom], 4.5), axis=1)
    assert np.all((d > 0.099) & (d < 0.151))
    m0 = sy.pos[:19]
    c = m0 - m0.mean(0)
    assert np.linalg.svd(c, compute_uv=False)[2] < 1e-12  # planar


def test_snapshot_batch():
    sizes = [S.config3_snapshot(k).n for k in range(64)]
    assert all(768 <= s <= 6144 for s in sizes) and len(set(sizes)) > 30
    sy = S.config3_snapshot(0)
    K = sy.settings["grid"][0]
    assert K >= sy.box[0, 0] / 0.08 and set(np.unique([K % 2, K % 3])) <= {0, 1}
    nmol = sy.n // 3
    assert sy.box[0, 0] == pytest.approx((nmol / 33.429) ** (1 / 3))


def test_prmtop_reader(tmp_path):
    text = """%VERSION  VERSION_STAMP = V0001.000
%FLAG POINTERS
%FORMAT(10I8)
       3       1
%FLAG POL_GAUSS_COVALENT_POINTERS_LIST
%FORMAT(10I8)
       2       1       1
%FLAG POL_GAUSS_COVALENT_ATOMS_LIST
%FORMAT(10I8)
       2       3       1       1
%FLAG POL_GAUSS_COVALENT_DIPOLES_LIST
%FORMAT(5E16.8)
 -1.91200000E-01 -1.91200000E-01  8.58770000E-02  8.58770000E-02
%FLAG POL_GAUSS_MONOPOLES_LIST
%FORMAT(5E16.8)
 -2.04060000E+00  1.02030000E+00  1.02030000E+00
%FLAG POL_GAUSS_RADII_LIST
%FORMAT(5E16.8)
  6.05150000E-01  5.36230000E-01  5.36230000E-01
%FLAG POL_GAUSS_POLARIZABILITY_LIST
%FORMAT(5E16.8)
  1.11765000E+00  3.29630000E-01  3.29630000E-01
"""
    p = tmp_path / "w.prmtop"
    p.write_text(text)
    par = prmtop.pgm_parameters(str(p))
    np.testing.assert_array_equal(par["cov_atom"], [0, 0, 1, 2])
    np.testing.assert_array_equal(par["cov_partner"], [1, 2, 0, 0])
    assert np.allclose(par["cov_strength"], [-0.01912, -0.01912, 0.0085877, 0.0085877])
    assert np.allclose(par["radius"], [0.060515, 0.053623, 0.053623])
    assert np.allclose(par["alpha"], [1.11765e-3, 0.32963e-3, 0.32963e-3])
    assert np.allclose(par["q"], [-2.0406, 1.0203, 1.0203])
    with pytest.raises(ValueError):
        p2 = tmp_path / "x.prmtop"
        p2.write_text("%FLAG POINTERS\n%FORMAT(10I8)\n       3\n")
        prmtop.pgm_parameters(str(p2))


def test_golden_fixture_parameters(golden):
    """pGM3P-25 water (docs/spec.md Sec. 2): q -2.0406/1.0203, R 0.60515/0.53623 A, alpha 1.11765/0.32963 A^3,
    O carries two covalent dipoles c = -0.19120 e A, each H one c = 0.085877 e A."""
    close = dict(rtol=1e-4, atol=0)  # the specification quotes five significant digits
    assert np.allclose(golden["q"][:3], [-2.0406, 1.0203, 1.0203], **close)
    assert np.allclose(golden["radius"][:3], [0.060515, 0.053623, 0.053623], **close)
    assert np.allclose(golden["alpha"][:3], [1.11765e-3, 0.32963e-3, 0.32963e-3], **close)
    assert np.allclose(golden["cov_strength"][:4], [-0.019120, -0.019120, 0.0085877, 0.0085877], **close)
    np.testing.assert_array_equal(golden["cov_atom"][:4], [0, 0, 1, 2])
    np.testing.assert_array_equal(golden["cov_partner"][:4], [1, 2, 0, 0])


def test_parameter_types_round_trip(golden):
    """Types group atoms with equal parameters; expanding the per-type values gives the per-atom ones back."""
    from pgmhip.prmtop import parameter_types

    params = {k: golden[k] for k in ("q", "radius", "alpha", "cov_atom", "cov_partner", "cov_strength")}
    ty = parameter_types(params)
    at, ct = ty["atom_type"], ty["cov_type"]
    assert len(ty["q"]) == 2  # water: O and H
    assert np.array_equal(ty["q"][at], params["q"])
    assert np.array_equal(ty["radius"][at], params["radius"])
    assert np.array_equal(ty["alpha"][at], params["alpha"])
    assert np.array_equal(ty["cov_strength"][ct], params["cov_strength"])
    assert np.array_equal(ty["cov_atom"], params["cov_atom"]) and np.array_equal(ty["cov_partner"], params["cov_partner"])
    # names split otherwise equal atoms
    names = np.where(np.arange(len(at)) % 2 == 0, "A", "B")
    ty2 = parameter_types(params, names)
    assert len(ty2["q"]) > len(ty["q"])
    assert np.array_equal(ty2["q"][ty2["atom_type"]], params["q"])
    with pytest.raises(ValueError):
        parameter_types(params, names[:-1])


def test_prmtop_field_layouts_and_errors(tmp_path):
    """%FORMAT specifications as AMBER writes them, comments between %FLAG and %FORMAT, and the checks of
    pgm_parameters (missing sections, counts that do not add up, partners out of range)."""
    L = prmtop.FieldLayout.parse
    assert L("(10I8)") == prmtop.FieldLayout(10, "i", 8)
    assert L("(5E16.8)") == prmtop.FieldLayout(5, "f", 16)
    assert L("(20a4)") == prmtop.FieldLayout(20, "a", 4)
    assert L("(i5)") == prmtop.FieldLayout(1, "i", 5)
    assert L("(3F12.6)  ") == prmtop.FieldLayout(3, "f", 12)
    for bad in ("10I8", "(10X8)", "(10I)", "()"):
        with pytest.raises(ValueError):
            L(bad)
    assert L("(4I3)").fields("  1 -2  3   ") == ["1", "-2", "3"]

    def write(name, counts, partners, strengths, nq=2):
        sec = lambda flag, fmt, body: f"%FLAG {flag}\n%COMMENT a comment\n%FORMAT({fmt})\n{body}\n"  # noqa: E731
        text = "%VERSION  VERSION_STAMP = V0001.000\n"
        text += sec("POL_GAUSS_COVALENT_POINTERS_LIST", "10I8", "".join(f"{c:8d}" for c in counts))
        text += sec("POL_GAUSS_COVALENT_ATOMS_LIST", "10I8", "".join(f"{c:8d}" for c in partners))
        text += sec("POL_GAUSS_COVALENT_DIPOLES_LIST", "5E16.8", "".join(f"{c:16.8E}" for c in strengths))
        for flag in ("POL_GAUSS_MONOPOLES_LIST", "POL_GAUSS_RADII_LIST", "POL_GAUSS_POLARIZABILITY_LIST"):
            text += sec(flag, "5E16.8", "".join(f"{0.5:16.8E}" for _ in range(nq)))
        p = tmp_path / name
        p.write_text(text)
        return str(p)

    par = prmtop.pgm_parameters(write("ok.prmtop", [1, 0], [2], [0.25]))
    np.testing.assert_array_equal(par["cov_atom"], [0])
    np.testing.assert_array_equal(par["cov_partner"], [1])
    assert par["cov_strength"][0] == 0.25 * 0.1 and par["alpha"][0] == 0.5 * 1e-3
    with pytest.raises(ValueError, match="add to"):
        prmtop.pgm_parameters(write("count.prmtop", [2, 0], [2], [0.25]))
    with pytest.raises(ValueError, match="outside"):
        prmtop.pgm_parameters(write("range.prmtop", [1, 0], [3], [0.25]))
    with pytest.raises(ValueError, match="entries for"):
        prmtop.pgm_parameters(write("short.prmtop", [1, 0, 0], [2], [0.25]))
    secs = prmtop.read_sections(write("all.prmtop", [1, 0], [2], [0.25]), ("POL_GAUSS_RADII_LIST", "NOT_THERE"))
    assert list(secs) == ["POL_GAUSS_RADII_LIST"] and secs["POL_GAUSS_RADII_LIST"].dtype == np.float64
