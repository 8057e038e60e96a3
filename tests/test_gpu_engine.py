"""Parity of the HIP engine (through the C ABI) with the FP64 oracle and the reference's golden vectors.

Tolerances: the FP64 path is held to the oracle at the level of FP64 round-off and solver tolerance (the
pGM-JAX agreement of the reference, docs/openmm.md:102-111, is 1e-12 rel energy and 2e-10 kJ/mol/nm); the
mixed path to its bound of 5e-5 relative-RMS force error and 1e-6 relative energy.
"""

import numpy as np
import pytest

import oracle as O
from conftest import golden_system, rel_rms, triclinic_waters
from pgmhip import engine as E
from pgmhip import systems as S

pytestmark = pytest.mark.gpu

MIXED_FORCE_RELRMS = 5e-5
MIXED_ENERGY_REL = 1e-6


def _energy(t):
    return E.total_energy(t)


def _oracle_ref(sy, **kw):
    o = O.Oracle(sy, **kw)
    e, f = o.compute(sy.pos, sy.box)
    return o, e, f


@pytest.mark.parametrize("level", ["q", "qp", "qi", "qpi"])
def test_golden_fp64(golden, level):
    sy = golden_system(golden, level)
    eng = E.Engine.for_system(sy, precision="double")
    t, f, _ = eng.compute(sy.pos, sy.box)
    er = float(golden[f"{level}_energy"])
    assert abs(_energy(t) / er - 1) < 1e-11
    assert np.abs(f - golden[f"{level}_forces"]).max() < 1e-8
    assert np.abs(eng.induced_dipoles() - golden[f"{level}_mu"]).max() < 1e-12
    assert eng.iterations() == int(golden[f"{level}_iterations"])


@pytest.mark.parametrize("level", ["q", "qpi"])
def test_golden_mixed(golden, level):
    sy = golden_system(golden, level)
    eng = E.Engine.for_system(sy, precision="mixed")
    t, f, _ = eng.compute(sy.pos, sy.box)
    er = float(golden[f"{level}_energy"])
    assert abs(_energy(t) / er - 1) < MIXED_ENERGY_REL
    assert rel_rms(f, golden[f"{level}_forces"]) < MIXED_FORCE_RELRMS
    assert eng.iterations() == int(golden[f"{level}_iterations"])


@pytest.mark.parametrize("maker", [S.config0, S.config1], ids=["water512", "ring8550"])
def test_workload_fp64_and_mixed_vs_oracle(maker):
    """The workload's own settings (tau, predictor, peek): terms, forces, dipoles, strain, iterations."""
    sy = maker()
    o, eo, fo = _oracle_ref(sy)
    to = o.terms()
    eng = E.Engine.for_system(sy, precision="double")
    t, f, Sx = eng.compute(sy.pos, sy.box, strain=True)
    for k in to:
        assert abs(t[k] - to[k]) <= 1e-10 * abs(eo), k
    assert np.abs(f - fo).max() <= 1e-9 * np.abs(fo).max()
    assert eng.iterations() == o.iterations()
    np.testing.assert_allclose(eng.induced_dipoles(), o.induced(), rtol=0, atol=1e-12)
    np.testing.assert_allclose(eng.permanent_dipoles(), o.permanent(), rtol=0, atol=1e-14)
    np.testing.assert_allclose(Sx, o.strain(), rtol=0, atol=1e-9 * np.abs(o.strain()).max())
    mix = E.Engine.for_system(sy, precision="mixed")
    tm, fm, Sm = mix.compute(sy.pos, sy.box, strain=True)
    assert abs(_energy(tm) / eo - 1) < MIXED_ENERGY_REL
    assert rel_rms(fm, fo) < MIXED_FORCE_RELRMS
    assert np.abs(Sm - o.strain()).max() < 1e-5 * np.abs(o.strain()).max()


def test_triclinic_tight():
    """The reference's finite-difference system (triclinic box, tau 1e-13), FP64."""
    sy = triclinic_waters(24, 7)
    o, eo, fo = _oracle_ref(sy)
    eng = E.Engine.for_system(sy, precision="double")
    t, f, Sx = eng.compute(sy.pos, sy.box, strain=True)
    assert abs(_energy(t) - eo) <= 1e-12 * abs(eo)
    assert np.abs(f - fo).max() <= 1e-9 * np.abs(fo).max()
    np.testing.assert_allclose(Sx, o.strain(), rtol=0, atol=1e-9 * np.abs(o.strain()).max())


def test_predictor_trajectory_matches_oracle():
    """mu4 predictor + peek over a trajectory with same-step repeats: the GPU engine's new_step flag against
    the oracle's history keys (core/pgm_core.cpp:307-328, 386-397; core/cuda/pgm_cuda.cu:2868-2884)."""
    sy = S.config0()
    traj = S.rigid_moves(sy, 3, 0.002, 6)
    seq = [(traj[0], 0, True), (traj[1], 1, True), (traj[1], 1, False), (traj[2], 2, True), (traj[3], 3, True),
           (traj[4], 4, True), (traj[5], 5, True), (traj[5], 5, False)]  # fmt: skip
    o = O.Oracle(sy)
    eng = E.Engine.for_system(sy, precision="double")
    for pos, key, new in seq:
        o.set_history_key(key)
        eo, fo = o.compute(pos, sy.box)
        t, f, _ = eng.compute(pos, sy.box, new_step=new)
        assert eng.iterations() == o.iterations()
        assert abs(_energy(t) - eo) <= 1e-10 * abs(eo)
        assert np.abs(f - fo).max() <= 1e-8 * np.abs(fo).max()


def test_history_closed_form():
    """With 4 loaded history sets, no iterations and no peek the dipoles are exactly 4 h0 - 6 h1 + 4 h2 - h3."""
    sy = S.config0()
    eng = E.Engine.for_system(sy, precision="double", max_iter=0, peek=0.0)
    rng = np.random.default_rng(5)
    H = rng.normal(0, 1e-3, (4, sy.n, 3))
    H[:, sy.alpha <= 1e-9] = 0
    eng.set_history(H)
    eng.compute(sy.pos, sy.box)
    np.testing.assert_allclose(eng.induced_dipoles(), 4 * H[0] - 6 * H[1] + 4 * H[2] - H[3], rtol=0, atol=1e-15)
    # mu3 with three sets
    eng2 = E.Engine.for_system(sy, precision="double", max_iter=0, peek=0.0, predictor=1)
    eng2.set_history(H[:3])
    eng2.compute(sy.pos, sy.box)
    np.testing.assert_allclose(eng2.induced_dipoles(), 3 * H[0] - 3 * H[1] + H[2], rtol=0, atol=1e-15)


def test_pair_list_candidates_and_rebuild():
    sy = S.config0()
    eng = E.Engine.for_system(sy, precision="double")
    eng.compute(sy.pos, sy.box)
    rl = 0.9 + 0.15
    L = sy.box[0, 0]
    d = sy.pos[:, None, :] - sy.pos[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d * d).sum(-1)
    np.fill_diagonal(r2, np.inf)
    assert eng.num_candidates() == int((r2 < rl * rl).sum())
    assert eng.list_builds() == 1
    small = sy.pos + 0.03  # a rigid shift below skin/2 per atom
    eng.compute(small, sy.box)
    assert eng.list_builds() == 1
    big = sy.pos.copy()
    big[5] += [0.09, 0.0, 0.0]  # one atom beyond skin/2
    t, f, _ = eng.compute(big, sy.box)
    assert eng.list_builds() == 2
    o = O.Oracle(sy)
    eo, fo = o.compute(big, sy.box)
    assert abs(_energy(t) - eo) <= 1e-9 * abs(eo)


def test_compute_f32_and_device_paths():
    torch = pytest.importorskip("torch")
    sy = S.config0()
    eng = E.Engine.for_system(sy, precision="mixed")
    p32 = sy.pos.astype(np.float32)
    t32, f32 = eng.compute_f32(p32, sy.box)
    ref = E.Engine.for_system(sy, precision="mixed")
    tr, fr, _ = ref.compute(p32.astype(np.float64), sy.box)
    # two mixed evaluations differ by the order of the FP32 atomic grid accumulation (~1e-7)
    assert abs(_energy(t32) - _energy(tr)) < 1e-8 * abs(_energy(tr))
    assert rel_rms(f32, fr) < 1e-6
    # float4 positions and fixed-point forces on the device (the OpenMM CUDA wrapper's contract)
    dev = torch.device("cuda:0")
    posq = torch.zeros((sy.n, 4), dtype=torch.float32, device=dev)
    posq[:, :3] = torch.from_numpy(p32).to(dev)
    stride = sy.n + 7
    fixed = torch.zeros(3 * stride, dtype=torch.int64, device=dev)
    eng3 = E.Engine.for_system(sy, precision="mixed")
    scale = 2.0**32
    td = eng3.compute_device(posq.data_ptr(), sy.box, True, fixed.data_ptr(), stride, scale,
                             torch.cuda.current_stream().cuda_stream)  # fmt: skip
    torch.cuda.synchronize()
    fx = fixed.cpu().numpy().astype(np.float64) / scale
    fdev = np.stack([fx[:sy.n], fx[stride:stride + sy.n], fx[2 * stride:2 * stride + sy.n]], 1)
    assert abs(_energy(td) - _energy(tr)) < 1e-8 * abs(_energy(tr))
    assert rel_rms(fdev, fr) < 1e-6
    # FP64 device arrays
    eng4 = E.Engine.for_system(sy, precision="double")
    pd = torch.from_numpy(sy.pos.copy()).to(dev)
    fd = torch.zeros((sy.n, 3), dtype=torch.float64, device=dev)
    t4 = eng4.compute_device_f64(pd.data_ptr(), sy.box, True, fd.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    o = O.Oracle(sy)
    eo, fo = o.compute(sy.pos, sy.box)
    assert abs(_energy(t4) - eo) <= 1e-10 * abs(eo)
    assert np.abs(fd.cpu().numpy() - fo).max() <= 1e-9 * np.abs(fo).max()


def test_launch_finish_and_wait():
    """launchDevice + finishDevice, and launchDevice + waitDevice on another stream (core/cuda/pgm_cuda.h:95-111):
    the forces of compute_device, read on the waiting stream without a host wait."""
    torch = pytest.importorskip("torch")
    sy = S.config0()
    dev = torch.device("cuda:0")
    posq = torch.zeros((sy.n, 4), dtype=torch.float32, device=dev)
    posq[:, :3] = torch.from_numpy(sy.pos.astype(np.float32)).to(dev)
    stride, scale = sy.n, 2.0**32
    ref = E.Engine.for_system(sy, precision="mixed")
    f0 = torch.zeros(3 * stride, dtype=torch.int64, device=dev)
    t0 = ref.compute_device(posq.data_ptr(), sy.box, True, f0.data_ptr(), stride, scale,
                            torch.cuda.current_stream().cuda_stream)  # fmt: skip
    torch.cuda.synchronize()
    eng = E.Engine.for_system(sy, precision="mixed")
    f1 = torch.zeros_like(f0)
    eng.launch_device(posq.data_ptr(), sy.box, True, f1.data_ptr(), stride, scale,
                      torch.cuda.current_stream().cuda_stream)  # fmt: skip
    t1, valid = eng.finish_device()
    torch.cuda.synchronize()
    assert valid and abs(_energy(t1) / _energy(t0) - 1) < 1e-8
    side = torch.cuda.Stream()
    f2 = torch.zeros_like(f0)
    eng2 = E.Engine.for_system(sy, precision="mixed")
    eng2.launch_device(posq.data_ptr(), sy.box, True, f2.data_ptr(), stride, scale,
                       torch.cuda.current_stream().cuda_stream)  # fmt: skip
    eng2.wait_device(side.cuda_stream)
    with torch.cuda.stream(side):
        f2d = f2.to(torch.float64) / scale  # ordered after the forces on the waiting stream
    side.synchronize()
    ref_f = f0.to(torch.float64) / scale
    for f in (f1.to(torch.float64) / scale, f2d):
        assert float(((f - ref_f) ** 2).sum().sqrt() / (ref_f**2).sum().sqrt()) < 1e-6


def test_errors_are_reported():
    sy = S.config0()
    with pytest.raises(E.PgmError):
        E.Engine(order=3)
    with pytest.raises(E.PgmError):
        E.Engine(grid=(32, 32, 11 * 13))  # prime factor 11 is not supported by the in-house FFT
    eng = E.Engine.for_system(sy, precision="double")
    with pytest.raises(E.PgmError):
        eng.compute(sy.pos, np.array([[2.48, 0.1, 0], [0, 2.48, 0], [0, 0, 2.48]]))  # not reduced
    with pytest.raises(E.PgmError):
        eng.set_covalent([0], [sy.n], [0.1])
    small = E.Engine.for_system(sy, precision="double")
    with pytest.raises(E.PgmError):
        small.compute(sy.pos, np.diag([1.5, 1.5, 1.5]))  # cutoff above half the box
    raw = E.Engine(precision="double", grid=(32, 32, 32))
    with pytest.raises(E.PgmError):
        raw.compute(sy.pos, sy.box)  # before set_atoms


def test_large_water_properties():
    """96,000 atoms (3x liquid density): mixed vs FP64 on the GPU, translation invariance, and the FP64 path
    against the oracle."""
    sy = S.config2()
    d = E.Engine.for_system(sy, precision="double")
    td, fd, _ = d.compute(sy.pos, sy.box)
    m = E.Engine.for_system(sy, precision="mixed")
    tm, fm, _ = m.compute(sy.pos, sy.box)
    assert m.iterations() == d.iterations()
    assert rel_rms(fm, fd) < MIXED_FORCE_RELRMS
    assert abs(_energy(tm) / _energy(td) - 1) < MIXED_ENERGY_REL
    # a shift by whole grid spacings leaves the PME discretization and every pair distance unchanged
    h = sy.box[0, 0] / sy.settings["grid"][0]
    shift = h * np.array([3.0, -2.0, 5.0])
    d2 = E.Engine.for_system(sy, precision="double")
    ts, fs, _ = d2.compute(sy.pos + shift, sy.box)
    assert d2.iterations() == d.iterations()
    assert abs(_energy(ts) / _energy(td) - 1) < 1e-11
    assert rel_rms(fs, fd) < 1e-9
    o = O.Oracle(sy)
    eo, fo = o.compute(sy.pos, sy.box)
    assert o.iterations() == d.iterations()
    assert abs(_energy(td) / eo - 1) < 1e-11
    assert rel_rms(fd, fo) < 1e-10
    assert rel_rms(fm, fo) < MIXED_FORCE_RELRMS


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("maker", [S.config0, S.config1], ids=["water512", "ring8550"])
def test_evaluations_are_bitwise_reproducible(precision, maker):
    """The PME grids accumulate in 64-bit fixed point (integer atomics commute) and every other sum runs in a
    fixed order, so evaluations are reproducible bit for bit: fresh engines, and the PME beside the pair passes
    or in line with them (timed phases, or pgm_set_pme_overlap 0), give the same energy terms, forces, strain and
    induced dipoles."""
    sy = maker()
    out = []
    for timed, beside in ((False, True), (False, True), (True, True), (False, False)):
        eng = E.Engine.for_system(sy, precision=precision)
        eng.set_timing(timed)
        eng.set_pme_overlap(beside)
        assert eng.pme_overlap() == beside
        t, f, S3 = eng.compute(sy.pos, sy.box, strain=True)
        out.append((t, f, S3, eng.induced_dipoles(), eng.iterations()))
    t0, f0, S0, m0, i0 = out[0]
    for t, f, S3, m, i in out[1:]:
        assert t == t0 and i == i0
        np.testing.assert_array_equal(f, f0)
        np.testing.assert_array_equal(S3, S0)
        np.testing.assert_array_equal(m, m0)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_pme_beside_pair_passes_same_results(precision):
    """The PME solves run on a second stream beside the pair passes, or on the engine stream while the PME phases
    are timed. The kernels and their inputs are the same either way. The evaluations are bitwise equal (fixed-point
    grids); the adjoint's tangent grids accumulate with floating-point atomics, whose order varies from run to
    run, which moves FP64 gradients by ulps and FP32 ones by their rounding. A missing fork/join dependency would
    read half-written fields: errors of order one, far above these bounds (FP64 1e-12, mixed 1e-6 relative)."""
    sy = S.config0()
    sy.settings.update(tol=1e-8, max_iter=200, predictor=0, peek=0.0)
    g = np.random.default_rng(2).normal(0.0, 1e-3, (sy.n, 3))
    out = []
    for timed in (False, True, False):
        eng = E.Engine.for_system(sy, precision=precision)
        eng.set_timing(timed)
        t, f, S3 = eng.compute(sy.pos, sy.box, strain=True)
        gr = eng.backward(sy.pos, sy.box, g, 0.5)
        out.append((t, f, S3, gr))
    tol = 1e-12 if precision == "double" else 1e-6
    t0, f0, S0, g0 = out[0]
    for t, f, S3, gr in out[1:]:
        assert t == t0
        np.testing.assert_array_equal(f, f0)
        np.testing.assert_array_equal(S3, S0)
        for k in gr:
            assert np.abs(gr[k] - g0[k]).max() <= tol * np.abs(g0[k]).max(), k


def test_autotune_picks_pme_overlap():
    """The autotuner times a CG product with the PME beside the field pass and after it and keeps the faster;
    the evaluations after it give the same results with either choice, and those of the oracle."""
    sy = S.config1()
    sy.settings.update(predictor=0)   # every evaluation from the alpha E guess
    eng = E.Engine.for_system(sy, precision="double")
    eng.compute(sy.pos, sy.box)
    assert eng.pme_overlap_times() == (0.0, 0.0)
    eng.autotune(2)
    after, beside = eng.pme_overlap_times()
    assert after > 0 and beside > 0
    assert eng.pme_overlap() == (beside <= after)
    t0, f0, _ = eng.compute(sy.pos, sy.box)
    eng.set_pme_overlap(not eng.pme_overlap())
    t1, f1, _ = eng.compute(sy.pos, sy.box)
    assert t1 == t0
    np.testing.assert_array_equal(f1, f0)
    eo, fo = O.Oracle(sy).compute(sy.pos, sy.box)
    assert abs(E.total_energy(t0) / eo - 1) < 1e-10
    assert rel_rms(f0, fo) < 1e-9


def test_engines_sharing_the_pme_stream():
    """Every engine on a device runs its PME solves on one shared stream (a hardware queue of its own).  Two
    engines of different systems and precisions, evaluating alternately on the caller's stream, give the same
    bits as each engine evaluating alone (every evaluation from the alpha E guess); the stream outlives the
    first engine's destruction, and the engines made after the last one was destroyed make it again."""
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    systems = [(S.config0(), "double"), (S.config1(), "mixed")]
    for sy, _ in systems:
        sy.settings.update(predictor=0)
    alone = []
    for sy, prec in systems:
        eng = E.Engine.for_system(sy, precision=prec)
        t, f, _ = eng.compute(sy.pos, sy.box)
        alone.append((t, f))
        del eng
    engines = [E.Engine.for_system(sy, precision=prec) for sy, prec in systems]
    pos = [torch.from_numpy(sy.pos.copy()).to(dev) for sy, _ in systems]
    forces = [torch.zeros((sy.n, 3), dtype=torch.float64, device=dev) for sy, _ in systems]
    terms = [None, None]
    for _ in range(3):
        for k, (sy, _) in enumerate(systems):
            forces[k].zero_()
            terms[k] = engines[k].compute_device_f64(pos[k].data_ptr(), sy.box, True, forces[k].data_ptr(), stream)
    torch.cuda.synchronize()
    for k in range(2):
        assert terms[k] == alone[k][0]
        np.testing.assert_array_equal(forces[k].cpu().numpy(), alone[k][1])
    engines[0] = None   # destroyed; the other engine keeps the shared stream
    sy, _ = systems[1]
    forces[1].zero_()
    t = engines[1].compute_device_f64(pos[1].data_ptr(), sy.box, True, forces[1].data_ptr(), stream)
    torch.cuda.synchronize()
    assert t == alone[1][0]
    np.testing.assert_array_equal(forces[1].cpu().numpy(), alone[1][1])
