"""Pins of the oracle's CG preconditioners (oracle/pgm_oracle.cpp: buildPreconditioner, precondition; options
of the reference's CUDA engine, core/cuda/pgm_cuda.cu:1526-1588,1753-1760; pgm_core itself is Jacobi).

* The short-range preconditioner is M = D + D T_s D, D = diag(alpha), T_s the direct-space dipole field tensor
  restricted to the pairs within the radius.  Without a box the field is the pair sum alone, so the tensor T
  assembled column by column from the oracle's field of unit dipoles (orc_field, pinned to the reference and to
  the golden vectors) is exactly the direct-space tensor: M r must equal (D + D T D) r with every pair inside
  the radius, and (D + D (T masked to the blocks within the radius) D) r with fewer; atoms without
  polarizability get zero rows and columns.
* More than 48 short-range pairs at any atom turn it off everywhere: M = D.
* Jacobi with the self term: z = alpha/(1 - alpha c_self) r, c_self = 4 beta^3/(3 sqrt(pi)).
* Preconditioned CG converges to the dense solution of (alpha^-1 - T) mu = E_perm, in fewer iterations than
  Jacobi at a tight tolerance; the tolerance stays on max |alpha r|.
"""

import numpy as np
import pytest

import oracle as O
from conftest import triclinic_waters


def _dense_T(o, n):
    zq = np.zeros(n)
    T = np.zeros((3 * n, 3 * n))
    for k in range(3 * n):
        e = np.zeros(3 * n)
        e[k] = 1.0
        T[:, k] = o.field(zq, e.reshape(n, 3)).reshape(-1)
    return T


def _cluster(nmol, seed, zero_alpha=()):
    sy = triclinic_waters(nmol, seed, periodic=False)
    sy.alpha = sy.alpha.copy()
    sy.alpha[list(zero_alpha)] = 0.0
    return sy


@pytest.mark.parametrize("zero_alpha", [(), (1, 5, 9)])
def test_short_range_m_is_d_plus_dtd(zero_alpha):
    sy = _cluster(12, 3, zero_alpha)  # 36 atoms: no atom can have more than 48 partners
    n = sy.n
    o = O.Oracle(sy)
    o.set_geometry(sy.pos, None)
    T = _dense_T(o, n)
    D = np.diag(np.repeat(sy.alpha, 3))
    r = np.random.default_rng(1).standard_normal((n, 3))
    o.set_preconditioner(100.0)
    z = o.precondition(r)
    assert o.preconditioner_on()
    M = D + D @ T @ D
    np.testing.assert_allclose(M, M.T, atol=1e-12 * np.abs(M).max())
    np.testing.assert_allclose(z.reshape(-1), M @ r.reshape(-1), rtol=0, atol=1e-12 * np.abs(z).max())
    # a radius that keeps some of the pairs: the tensor's blocks within it
    rad = 0.35
    dist = np.linalg.norm(sy.pos[:, None, :] - sy.pos[None, :, :], axis=-1)
    assert 0 < (dist < rad).sum() - n < n * (n - 1)
    mask = np.kron((dist < rad).astype(float), np.ones((3, 3)))
    o.set_preconditioner(rad)
    z = o.precondition(r)
    Mr = (D + D @ (T * mask) @ D) @ r.reshape(-1)
    np.testing.assert_allclose(z.reshape(-1), Mr, rtol=0, atol=1e-12 * np.abs(z).max())


def test_cap_turns_it_off():
    sy = _cluster(20, 4)  # 60 atoms, all within the radius: 59 partners each
    o = O.Oracle(sy)
    o.set_geometry(sy.pos, None)
    o.set_preconditioner(100.0)
    r = np.random.default_rng(2).standard_normal((sy.n, 3))
    z = o.precondition(r)
    assert not o.preconditioner_on()
    np.testing.assert_array_equal(z, sy.alpha[:, None] * r)


def test_jacobi_self_term():
    sy = triclinic_waters(8, 5)
    o = O.Oracle(sy)
    o.set_geometry(sy.pos, sy.box)
    o.set_preconditioner(0.0, jacobi_self=True)
    r = np.random.default_rng(3).standard_normal((sy.n, 3))
    cs = 4 * sy.settings["beta"] ** 3 / (3 * np.sqrt(np.pi))
    np.testing.assert_allclose(o.precondition(r), (sy.alpha / (1 - sy.alpha * cs))[:, None] * r, rtol=1e-15)


def test_radius_checked():
    sy = triclinic_waters(8, 5)
    o = O.Oracle(sy)
    with pytest.raises(RuntimeError, match="radius"):
        o.set_preconditioner(sy.settings["cutoff"] + 0.1)
    with pytest.raises(RuntimeError, match="radius"):
        o.set_preconditioner(-0.1)


@pytest.mark.parametrize("radius, jacobi_self", [(0.3, False), (0.0, True)])
def test_preconditioned_cg_solves_the_same_system(radius, jacobi_self):
    """Periodic waters: the preconditioned CG's dipoles equal the dense solve; at tau = 1e-12 the short-range
    preconditioner needs fewer iterations than Jacobi."""
    sy = triclinic_waters(8, 5)
    n = sy.n
    o = O.Oracle(sy, tol=1e-12, max_iter=5000, peek=0.0, predictor=0)
    o.compute(sy.pos, sy.box, forces=False)
    jacobi_its = o.iterations()
    T = _dense_T(o, n)
    A = np.diag(np.repeat(1.0 / sy.alpha, 3)) - T
    b = o.field(sy.q, o.permanent()).reshape(-1)
    mu_dense = np.linalg.solve(A, b)
    p = O.Oracle(sy, tol=1e-12, max_iter=5000, peek=0.0, predictor=0)
    p.set_preconditioner(radius, jacobi_self=jacobi_self)
    p.compute(sy.pos, sy.box, forces=False)
    assert p.preconditioner_on() == (radius > 0)
    np.testing.assert_allclose(p.induced().reshape(-1), mu_dense, rtol=0, atol=1e-9 * np.abs(mu_dense).max())
    assert p.residual() <= 1e-12
    print(f"radius {radius} self {jacobi_self}: {p.iterations()} iterations (Jacobi {jacobi_its})")
    if radius > 0:
        assert p.iterations() < jacobi_its
