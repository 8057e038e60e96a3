"""The hand-written adjoint (pgm_backward) against central finite differences of the FP64 loss.

The reference has no parameter derivatives, so only finite differences pin these gradients: of the oracle's
FP64 loss (an independent CPU evaluation) and of the engine's own FP64 loss.  Loss L = wE E + sum g . F with a
seeded cotangent g; per-type parameters are perturbed by h = 1e-5 relative (central differences), the solves
converged to tau = 1e-12 so that the solver's own error stays far below the differences.  Bound: 1e-5
relative per parameter (with a floor of 1e-7 of the largest gradient for parameters whose gradient is nearly
zero).
"""

import dataclasses

import numpy as np
import pytest

import oracle as O
from conftest import triclinic_waters
from pgmhip import engine as E
from pgmhip import systems as S
from pgmhip.prmtop import parameter_types

pytestmark = pytest.mark.gpu

FD_REL = 1e-5
TIGHT = dict(tol=1e-12, max_iter=2000, predictor=0, peek=0.0)


def _types(sy):
    return parameter_types(
        {"q": sy.q, "radius": sy.radius, "alpha": sy.alpha, "cov_atom": sy.cov_atom, "cov_partner": sy.cov_partner,
         "cov_strength": sy.cov_strength}
    )  # fmt: skip


def _theta(ty):
    return np.concatenate([ty["q"], ty["radius"], ty["alpha"], ty["cov_strength"]])


def _expand(sy, ty, theta):
    nt, nc = len(ty["q"]), len(ty["cov_strength"])
    tq, tr, ta, tc = theta[:nt], theta[nt : 2 * nt], theta[2 * nt : 3 * nt], theta[3 * nt : 3 * nt + nc]
    at = ty["atom_type"]
    return dataclasses.replace(sy, q=tq[at], radius=tr[at], alpha=ta[at], cov_strength=tc[ty["cov_type"]])


def _per_type(ty, grads):
    nt, nc = len(ty["q"]), len(ty["cov_strength"])
    at = ty["atom_type"]
    out = [np.bincount(at, grads[k], minlength=nt) for k in ("q", "radius", "alpha")]
    out.append(np.bincount(ty["cov_type"], grads["cov_strength"], minlength=nc))
    return np.concatenate(out)


def _loss_oracle(sy, g, wE):
    o = O.Oracle(sy)
    e, f = o.compute(sy.pos, sy.box)
    return wE * e + float((g * f).sum())


def _loss_engine(eng, sy, g, wE):
    eng.set_atoms(sy.q, sy.radius, sy.alpha)
    eng.set_covalent(sy.cov_atom, sy.cov_partner, sy.cov_strength)
    t, f, _ = eng.compute(sy.pos, sy.box, new_step=False)
    return wE * E.total_energy(t) + float((g * f).sum())


def _fd(loss, sy, ty, theta, h=1e-5):
    out = np.zeros_like(theta)
    for k in range(len(theta)):
        d = h * abs(theta[k]) if theta[k] != 0 else h
        tp, tm = theta.copy(), theta.copy()
        tp[k] += d
        tm[k] -= d
        out[k] = (loss(_expand(sy, ty, tp)) - loss(_expand(sy, ty, tm))) / (2 * d)
    return out


def _compare(ga, gf):
    floor = 1e-7 * np.abs(gf).max()
    rel = np.abs(ga - gf) / np.maximum(np.abs(gf), floor)
    return rel


def _cotangent(sy, seed=5, scale=1e-3):
    return np.random.default_rng(seed).normal(0.0, scale, (sy.n, 3))


def test_energy_gradient_matches_fd_per_atom():
    """dE/dtheta per atom and per covalent entry (no cotangent): the parameter derivatives at fixed dipoles."""
    sy = triclinic_waters(60, 3)
    eng = E.Engine.for_system(sy, precision="double")
    eng.compute(sy.pos, sy.box)
    gr = eng.backward(sy.pos, sy.box, None, 1.0)
    assert eng.adjoint_iterations() == 0
    zero = np.zeros((sy.n, 3))
    h = 1e-5
    for name, idx in (("q", [0, 1, 2, 100]), ("radius", [0, 1, 100]), ("alpha", [0, 2, 101]),
                      ("cov_strength", [0, 1, 7])):  # fmt: skip
        for i in idx:
            vals = {}
            for sgn in (1, -1):
                s2 = dataclasses.replace(sy, **{name: getattr(sy, name).copy()})
                arr = getattr(s2, name)
                d = h * abs(arr[i])
                arr[i] += sgn * d
                vals[sgn] = _loss_engine(eng, s2, zero, 1.0)
            fd = (vals[1] - vals[-1]) / (2 * d)
            assert abs(gr[name][i] - fd) <= FD_REL * abs(fd), (name, i, gr[name][i], fd)


def test_force_loss_gradient_matches_oracle_fd():
    """Triclinic box, covalent dipoles on O and H, every term active: the adjoint against central differences
    of the oracle's FP64 loss (an evaluation independent of the HIP code)."""
    sy = triclinic_waters(60, 3)
    ty = _types(sy)
    g = _cotangent(sy)
    wE = 0.3
    eng = E.Engine.for_system(sy, precision="double")
    eng.compute(sy.pos, sy.box)
    ga = _per_type(ty, eng.backward(sy.pos, sy.box, g, wE))
    assert eng.adjoint_iterations() > 0
    gf = _fd(lambda s: _loss_oracle(s, g, wE), sy, ty, _theta(ty))
    rel = _compare(ga, gf)
    assert rel.max() < FD_REL, (rel, ga, gf)


def test_snapshot_force_matching_gradient_matches_fd():
    """A snapshot of the uneven batch (258 waters, 1 g/cm^3), the force-matching loss's cotangent
    g = 2 (F - F_ref)/n with F_ref at theta x 1.05: adjoint against central differences of the engine's FP64 loss."""
    sy = S.config3_snapshot(41)
    sy.settings.update(TIGHT)
    ty = _types(sy)
    theta = _theta(ty)
    eng = E.Engine.for_system(sy, precision="double")
    ref = _expand(sy, ty, 1.05 * theta)
    eng.set_atoms(ref.q, ref.radius, ref.alpha)
    eng.set_covalent(ref.cov_atom, ref.cov_partner, ref.cov_strength)
    _, fref, _ = eng.compute(sy.pos, sy.box)
    eng.set_atoms(sy.q, sy.radius, sy.alpha)
    eng.set_covalent(sy.cov_atom, sy.cov_partner, sy.cov_strength)
    _, f, _ = eng.compute(sy.pos, sy.box)
    g = 2.0 * (f - fref) / sy.n
    ga = _per_type(ty, eng.backward(sy.pos, sy.box, g, 0.0))
    fd_eng = E.Engine.for_system(sy, precision="double")
    gf = _fd(lambda s: _loss_engine(fd_eng, s, g, 0.0), sy, ty, theta)
    rel = _compare(ga, gf)
    assert rel.max() < FD_REL, (rel, ga, gf)


def test_mixed_gradient_matches_fp64():
    """The mixed-precision adjoint against the FP64 one on the same snapshot (norm-relative)."""
    sy = S.config3_snapshot(41)
    sy.settings.update(tol=1e-10, max_iter=2000, predictor=0, peek=0.0)
    ty = _types(sy)
    g = _cotangent(sy, seed=7, scale=1.0)
    out = {}
    for prec in ("double", "mixed"):
        eng = E.Engine.for_system(sy, precision=prec)
        eng.compute(sy.pos, sy.box)
        out[prec] = _per_type(ty, eng.backward(sy.pos, sy.box, g, 0.5))
    err = np.abs(out["mixed"] - out["double"]) / np.abs(out["double"]).max()
    print("mixed vs FP64 per-type gradient, relative to the largest:", err)
    assert err.max() < 1e-4


def test_nonpolarizable_without_covalent_matches_oracle_fd():
    """No induction and no covalent dipoles: the adjoint reduces to the directional derivative (no solve)."""
    sy = triclinic_waters(40, 9)
    sy.settings.update(polarizable=0)
    sy = dataclasses.replace(sy, cov_strength=np.zeros_like(sy.cov_strength))
    ty = _types(sy)
    g = _cotangent(sy, seed=11)
    eng = E.Engine.for_system(sy, precision="double")
    eng.compute(sy.pos, sy.box)
    ga = _per_type(ty, eng.backward(sy.pos, sy.box, g, 1.0))
    assert eng.adjoint_iterations() == 0
    theta = _theta(ty)
    gf = _fd(lambda s: _loss_oracle(s, g, 1.0), sy, ty, theta)
    keep = np.ones(len(theta), bool)
    nt = len(ty["q"])
    keep[2 * nt : 3 * nt] = False  # polarizabilities: no induction, gradient exactly zero
    assert np.all(ga[~keep] == 0)
    rel = _compare(ga[keep], gf[keep])
    assert rel.max() < FD_REL, (rel, ga, gf)


def test_torch_function_matches_backward():
    """PgmModel / PgmFunction: loss.backward() gives the engine's per-atom gradients reduced over the types."""
    import torch

    from pgmhip.adjoint import PgmModel, force_matching_loss

    sy = S.config3_snapshot(41)
    sy.settings.update(TIGHT)
    ty = _types(sy)
    eng = E.Engine.for_system(sy, precision="double")
    model = PgmModel.from_types(eng, ty, sy.box)
    dev = torch.device("cuda:0")
    params = [torch.tensor(ty[k], dtype=torch.float64, device=dev, requires_grad=True)
              for k in ("q", "radius", "alpha", "cov_strength")]  # fmt: skip
    pos = torch.tensor(sy.pos, dtype=torch.float64, device=dev)
    fref = torch.zeros_like(pos)
    energy, _ = model(*params, pos)
    loss = force_matching_loss([model], params, [pos], [fref]) + 1e-3 * energy
    loss.backward()
    got = np.concatenate([p.grad.cpu().numpy() for p in params])
    # the first call contributes through dL/dE = 1e-3 only, the loss's own call through dL/dF = 2 F / n only
    ref_eng = E.Engine.for_system(sy, precision="double")
    _, f, _ = ref_eng.compute(sy.pos, sy.box)
    want = _per_type(ty, ref_eng.backward(sy.pos, sy.box, 2.0 * f / sy.n, 0.0))
    want += _per_type(ty, ref_eng.backward(sy.pos, sy.box, None, 1e-3))
    assert np.allclose(got, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())


def test_backward_errors():
    eng = E.Engine(precision="double")
    with pytest.raises(E.PgmError, match="backward before setAtoms"):
        eng.backward(np.zeros((0, 3)), np.eye(3), None, 1.0)


def test_backward_after_an_evaluation_at_the_same_positions():
    """A backward at the positions and box of the engine's last evaluation reuses that evaluation (no second
    forward solve): its gradients equal those of a backward that evaluates first (a fresh engine), also when
    the device positions come in another buffer with the same values; at other positions it evaluates again."""
    torch = pytest.importorskip("torch")
    sy = S.config0()
    sy.settings.update(TIGHT)
    g = np.random.default_rng(8).normal(0.0, 1e-3, (sy.n, 3))

    def fresh(pos):
        return E.Engine.for_system(sy, precision="double").backward(pos, sy.box, g, 0.3)

    def close(a, b, tol=1e-11):
        for k in a:
            assert np.abs(a[k] - b[k]).max() <= tol * np.abs(b[k]).max(), k

    ref = fresh(sy.pos)
    eng = E.Engine.for_system(sy, precision="double")
    eng.compute(sy.pos, sy.box)
    close(eng.backward(sy.pos, sy.box, g, 0.3), ref)
    # device positions in a buffer other than the evaluated one, same values
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    p1 = torch.from_numpy(sy.pos.copy()).to(dev)
    f1 = torch.zeros((sy.n, 3), dtype=torch.float64, device=dev)
    eng.compute_device_f64(p1.data_ptr(), sy.box, True, f1.data_ptr(), stream)
    p2 = p1.clone()
    del p1
    gd = torch.from_numpy(g).to(dev)
    outs = [torch.zeros(sy.n, dtype=torch.float64, device=dev) for _ in range(3)]
    gc = torch.zeros(len(sy.cov_atom), dtype=torch.float64, device=dev)
    eng.backward_device(p2.data_ptr(), sy.box, gd.data_ptr(), 0.3, *[t.data_ptr() for t in outs], gc.data_ptr(),
                        stream)  # fmt: skip
    torch.cuda.synchronize()
    close({"q": outs[0].cpu().numpy(), "radius": outs[1].cpu().numpy(), "alpha": outs[2].cpu().numpy(),
           "cov_strength": gc.cpu().numpy()}, ref)  # fmt: skip
    # other positions: evaluated there first
    moved = sy.pos + np.random.default_rng(9).normal(0.0, 2e-3, sy.pos.shape)
    eng.compute(sy.pos, sy.box)
    close(eng.backward(moved, sy.box, g, 0.3), fresh(moved), 1e-9)  # other guesses: the solver tolerance
