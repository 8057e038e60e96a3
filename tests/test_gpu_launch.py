"""The asynchronous device boundary: pgm_launch_device returns before the GPU work is done, pgm_finish_device
finds a stale pair list or a short solve late and asks for a repeat (valid = 0), and pgm_wait_device leaves
the completion to the next call, which counts such a find as late (core/cuda/pgm_cuda.h:95-111,
core/cuda/pgm_cuda.cu:2754-2788,2932-2959).

Every valid launch gives the forces and energies of pgm_compute_device over the same sequence of positions
(the same guesses and history entries).
"""

import numpy as np
import pytest

from pgmhip import engine as E
from pgmhip import systems as S

pytestmark = pytest.mark.gpu

SCALE = 2.0**32


@pytest.fixture(scope="module")
def torch():
    return pytest.importorskip("torch")


class Device:
    """float4 positions and a fixed-point force buffer on cuda:0, the OpenMM wrapper's contract."""

    def __init__(self, torch, n):
        self.t, self.n = torch, n
        self.stride = n + 3
        self.posq = torch.zeros((n, 4), dtype=torch.float32, device="cuda:0")
        self.fixed = torch.zeros(3 * self.stride, dtype=torch.int64, device="cuda:0")

    def load(self, pos):
        self.posq[:, :3] = self.t.from_numpy(pos.astype(np.float32)).to("cuda:0")
        self.fixed.zero_()

    @property
    def stream(self):
        return self.t.cuda.current_stream().cuda_stream

    def forces(self):
        self.t.cuda.synchronize()
        fx = self.fixed.cpu().numpy().astype(np.float64) / SCALE
        s = self.stride
        return np.stack([fx[: self.n], fx[s : s + self.n], fx[2 * s : 2 * s + self.n]], 1)

    def launch(self, eng, box, new_step=True):
        eng.launch_device(self.posq.data_ptr(), box, new_step, self.fixed.data_ptr(), self.stride, SCALE, self.stream)

    def compute(self, eng, box, new_step=True):
        t = eng.compute_device(self.posq.data_ptr(), box, new_step, self.fixed.data_ptr(), self.stride, SCALE,
                               self.stream)  # fmt: skip
        return t, self.forces()


def _same(t, f, tr, fr, precision):
    e, er = E.total_energy(t), E.total_energy(tr)
    if precision == "double":
        assert abs(e - er) <= 1e-10 * abs(er)
        assert np.abs(f - fr).max() <= 1e-8 * np.abs(fr).max()
    else:
        assert abs(e - er) <= 1e-7 * abs(er)
        assert E.rel_rms(f, fr) < 1e-6


def test_launch_returns_before_the_gpu_work_completes(torch):
    """The launch is ordered after the caller's stream and does not wait for it: with 50 ms of matrix products
    queued on the caller's stream first, the launch's work is still pending when the call returns."""
    sy = S.config1()
    dev = Device(torch, sy.n)
    eng = E.Engine.for_system(sy, precision="mixed")
    ref = E.Engine.for_system(sy, precision="mixed")
    traj = S.rigid_moves(sy, 4, 0.002, 3)
    a = torch.randn((4096, 4096), device="cuda:0")
    for k, pos in enumerate(traj):
        dev.load(pos)
        tr, fr = dev.compute(ref, sy.box)
        dev.load(pos)
        if k > 0:   # (the first launch of an engine waits on the host inside its solve)
            torch.cuda.synchronize()
            for _ in range(40):
                a = (a @ a) * 1e-3
        dev.launch(eng, sy.box)
        if k > 0:
            assert eng.launch_pending(), "the launch waited for the GPU"
        t, valid = eng.finish_device()
        assert valid and not eng.launch_pending()
        assert eng.iterations() == ref.iterations()
        _same(t, dev.forces(), tr, fr, "mixed")


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_stale_list_found_at_finish_then_repeat(torch, precision):
    """One atom moved beyond skin/2 between two launches: the host's prediction (from the last step) keeps the
    list, the device check finds it stale, finish says valid = 0 and has rebuilt it; the repeat is valid and
    equals pgm_compute_device at those positions after the same history."""
    sy = S.config0()
    dev = Device(torch, sy.n)
    pos0 = sy.pos
    pos1 = sy.pos.copy()
    pos1[5] += [0.09, 0.0, 0.0]
    eng = E.Engine.for_system(sy, precision=precision)
    ref = E.Engine.for_system(sy, precision=precision)
    dev.load(pos0)
    dev.compute(ref, sy.box)
    dev.load(pos0)
    dev.launch(eng, sy.box)
    t, valid = eng.finish_device()
    assert valid and eng.list_builds() == 1
    dev.load(pos1)
    tr, fr = dev.compute(ref, sy.box)
    dev.load(pos1)
    dev.launch(eng, sy.box)
    _, valid = eng.finish_device()
    assert not valid and eng.list_builds() == 2
    dev.load(pos1)
    dev.launch(eng, sy.box)
    t, valid = eng.finish_device()
    assert valid and eng.list_builds() == 2 and eng.late_list_checks() == 0
    assert eng.iterations() == ref.iterations()
    _same(t, dev.forces(), tr, fr, precision)


def test_stale_list_at_finish_with_phase_timing(torch):
    """The same stale list found at finish, with the list phase timed: the rebuild that finish enqueues is
    bracketed too, and its timing is read once it has run (it used to be read too early: 'device not ready')."""
    sy = S.config0()
    dev = Device(torch, sy.n)
    pos1 = sy.pos.copy()
    pos1[5] += [0.09, 0.0, 0.0]
    eng = E.Engine.for_system(sy, precision="mixed")
    eng.set_timing(True, phases=["list"])
    for pos, want in ((sy.pos, True), (pos1, False), (pos1, True)):
        dev.load(pos)
        dev.launch(eng, sy.box)
        _, valid = eng.finish_device()
        assert valid == want
    assert eng.list_builds() == 2
    ms, count = eng.timing()["list"]
    assert ms > 0 and count >= 2   # both builds bracketed (the first launch's and the one finish enqueued)


def test_short_solve_repeats(torch):
    """A launch enqueues the most iterations of the last 16 solves (one more if they varied); with a bad history
    (four sets of random dipoles, the same on both engines) the guess needs more, finish says valid = 0, and the
    repeat (which checks convergence on the host) is valid."""
    sy = S.config0()
    dev = Device(torch, sy.n)
    eng = E.Engine.for_system(sy, precision="double")
    ref = E.Engine.for_system(sy, precision="double")
    traj = S.rigid_moves(sy, 6, 0.001, 19)
    its = []
    for pos in traj[:18]:
        dev.load(pos)
        dev.compute(ref, sy.box)
        dev.load(pos)
        dev.launch(eng, sy.box)
        _, valid = eng.finish_device()
        assert valid
        its.append(eng.iterations())
    bad = np.random.default_rng(8).normal(0.0, 0.01, (4, sy.n, 3))   # |mu| is about 0.003 e nm
    eng.set_history(bad)
    ref.set_history(bad)
    dev.load(traj[18])
    tr, fr = dev.compute(ref, sy.box)
    assert ref.iterations() > max(its[-16:]) + 1, "the workload does not exercise a short solve"
    dev.load(traj[18])
    dev.launch(eng, sy.box)
    _, valid = eng.finish_device()
    assert not valid
    dev.load(traj[18])
    dev.launch(eng, sy.box)
    t, valid = eng.finish_device()
    assert valid and eng.iterations() == ref.iterations()
    _same(t, dev.forces(), tr, fr, "double")


def test_wait_without_finish_counts_late(torch):
    """launch + wait_device: the forces are read on another stream without a host wait; a stale list is found
    at the next call (counted late, the list rebuilt) and later evaluations are right."""
    sy = S.config0()
    dev = Device(torch, sy.n)
    eng = E.Engine.for_system(sy, precision="double")
    dev.load(sy.pos)
    dev.launch(eng, sy.box)
    side = torch.cuda.Stream()
    eng.wait_device(side.cuda_stream)
    with torch.cuda.stream(side):
        f0 = dev.fixed.to(torch.float64) / SCALE
    side.synchronize()
    assert float(f0.abs().max()) > 0
    moved = sy.pos.copy()
    moved[7] += [0.0, 0.09, 0.0]
    dev.load(moved)
    dev.launch(eng, sy.box)
    eng.wait_device(side.cuda_stream)
    side.synchronize()
    assert eng.late_list_checks() == 0   # not completed yet
    t, f, _ = eng.compute(moved, sy.box, new_step=False)   # any call completes the launch first
    assert eng.late_list_checks() == 1 and eng.list_builds() == 2
    fresh = E.Engine.for_system(sy, precision="double", tol=1e-12)
    tf, ff, _ = fresh.compute(moved, sy.box)
    assert abs(E.total_energy(t) / E.total_energy(tf) - 1) < 1e-6


def test_launch_predicts_rebuilds_along_a_trajectory(torch):
    """Atoms moving at constant velocity: every launch predicts from the last check's displacement and the last
    step's largest move (both measured on the device, reported by pgm_list_check_state) whether the list holds,
    and rebuilds before it goes stale, so no launch is repeated for a stale list.  The largest move is measured
    afresh at each launch (its word was once left uncleared, so it held whatever the device memory held and
    the prediction never fired)."""
    sy = S.config0()
    dev = Device(torch, sy.n)
    rng = np.random.default_rng(5)
    v = rng.normal(0.0, 1.0, sy.pos.shape)
    v /= np.sqrt((v * v).sum(1)).max()   # the fastest atom moves 0.01 nm per step: skin/2 in 7.5 steps
    step = 0.01
    eng = E.Engine.for_system(sy, precision="mixed")
    builds0, stale_repeats = None, 0
    prev = None
    for k in range(24):
        pos = sy.pos + k * step * v
        while True:
            dev.load(pos)
            b0 = eng.list_builds()
            dev.launch(eng, sy.box)
            _, valid = eng.finish_device()
            if valid:
                break
            stale_repeats += eng.list_builds() > b0
        st = eng.list_check_state()
        assert np.isfinite(st["last_move"]) and np.isfinite(st["displacement"])
        if prev is not None and k >= 2:
            p32, q32 = pos.astype(np.float32).astype(float), prev.astype(np.float32).astype(float)
            true_move = np.sqrt(((p32 - q32) ** 2).sum(1)).max()
            assert abs(st["last_move"] - true_move) < 1e-4, (k, st, true_move)
        prev = pos
        if k == 0:
            builds0 = eng.list_builds()
    assert eng.list_builds() - builds0 >= 2   # the list was rebuilt along the way
    assert stale_repeats == 0


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_sync_evaluation_with_cg_enqueued_ahead(torch, precision):
    """pgm_compute_device enqueues the CG ahead once the engine has solved before (no host wait between the solve
    and the final stage) and adds the forces only if the enqueued iterations sufficed.  Along a trajectory no
    solve falls short (the trajectory tests against the oracle cover these evaluations' results); after a bad
    history it does, the evaluation is redone from the same guess with its convergence checked on the host, the
    forces are added once, and the results equal those of a fresh engine (host-checked: it has not solved
    before) with the same history."""
    sy = S.config0()
    dev = Device(torch, sy.n)
    eng = E.Engine.for_system(sy, precision=precision)
    for pos in S.rigid_moves(sy, 6, 0.001, 8):
        dev.load(pos)
        dev.compute(eng, sy.box)
    assert eng.list_check_state()["sync_short_solves"] == 0
    bad = np.random.default_rng(8).normal(0.0, 0.01, (4, sy.n, 3))   # |mu| is about 0.003 e nm
    eng.set_history(bad)
    ref = E.Engine.for_system(sy, precision=precision)
    ref.set_history(bad)
    dev.load(sy.pos)
    t, f = dev.compute(eng, sy.box)
    dev.load(sy.pos)
    tr, fr = dev.compute(ref, sy.box)
    assert eng.list_check_state()["sync_short_solves"] == 1
    assert eng.iterations() == ref.iterations()
    _same(t, f, tr, fr, precision)
