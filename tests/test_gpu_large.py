"""The 999,000-atom water box (333,000 waters, 21.54 nm, 216^3 PME, order 6) on one GPU, against the FP64
oracle's evaluation of the same box at tau = 1e-10 (tests/golden/water999k_oracle.npz, written by
tests/golden/make_water999k.py: terms, iterations, strain, forces and dipoles of 3,000 seeded atoms, and
projections of all forces and dipoles on 8 seeded Gaussian vectors).

- FP64 path, tau 1e-10, no predictor, no peek: the oracle's iterations, energy terms to 1e-11 relative, forces of
  the sample to 1e-9 relative RMS, the projections (every atom) and sum |F|^2 alike, strain to 1e-9.
- Mixed path at the box's own settings (tau 1e-5, peek 0.65, guess alpha E; FP32 pair and PME arithmetic, FP64
  CG vectors and sums): the error against the oracle, which the run prints, within the 5e-5 relative-RMS force
  bound, both on the sample and as estimated from the projections.
- Device memory of each engine with its lists resident, and evaluations per second over MD-like steps (atoms
  moved by up to 1e-4 nm a step, the previous solution as the guess), printed.
"""

import os
import time

import numpy as np
import pytest
import torch

from pgmhip import engine as E
from pgmhip import systems as S

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TERMS = ("direct", "reciprocal", "self", "background", "polarization")


def _fixture():
    return np.load(os.path.join(HERE, "golden", "water999k_oracle.npz"))


def _projections(v):
    w = np.random.default_rng(999).standard_normal((8, v.shape[0], 3))
    return np.einsum("kna,na->k", w, v)


def _run(sy, precision, reps):
    free0 = torch.cuda.mem_get_info()[0]
    eng = E.Engine.for_system(sy, precision=precision)
    t, f, S3 = eng.compute(sy.pos, sy.box, strain=True)
    used = (free0 - torch.cuda.mem_get_info()[0]) / 2**30
    mu = eng.induced_dipoles()
    its = eng.iterations()
    # MD-like steps: every atom moved by up to 1e-4 nm a step, the previous solution as the guess
    rng = np.random.default_rng(7)
    steps = [sy.pos + rng.uniform(-1e-4, 1e-4, sy.pos.shape) * (k + 1) for k in range(reps)]
    t0 = time.perf_counter()
    for p in steps:
        eng.compute(p, sy.box, forces=True)
    rate = reps / (time.perf_counter() - t0) if reps else 0.0
    del eng
    return t, f, S3, mu, its, used, rate


def test_fp64_against_the_oracle():
    g = _fixture()
    sy = S.config4()
    sy.settings.update(tol=1e-10, max_iter=500, peek=0.0, predictor=0)
    t, f, S3, mu, its, used, _ = _run(sy, "double", 0)
    idx = g["sample"]
    print(f"FP64 999k: {its} iterations (oracle {int(g['iterations'])}), device memory {used:.1f} GiB")
    assert its == int(g["iterations"])
    for k, v in zip(TERMS, g["terms"]):
        assert abs(t[k] - v) <= 1e-11 * abs(v) + 1e-6, k
    fs = g["forces_sample"]
    assert np.sqrt(((f[idx] - fs) ** 2).sum() / (fs**2).sum()) < 1e-9
    assert np.abs(mu[idx] - g["induced_sample"]).max() < 1e-9 * np.abs(g["induced_sample"]).max()
    assert abs((f * f).sum() / g["forces_sumsq"] - 1) < 1e-9
    p, po = _projections(f), g["forces_proj"]
    assert np.sqrt(((p - po) ** 2).mean() / g["forces_sumsq"]) < 1e-9
    assert np.abs(S3 - g["strain"]).max() < 1e-9 * np.abs(g["strain"]).max()


def test_mixed_error_report():
    g = _fixture()
    sy = S.config4()
    t, f, _, _, its, used, rate = _run(sy, "mixed", 5)
    idx = g["sample"]
    fs = g["forces_sample"]
    rel_sample = np.sqrt(((f[idx] - fs) ** 2).sum() / (fs**2).sum())
    # E[(dF.w)^2] = |dF|^2: the projections estimate the error over every atom
    rel_proj = np.sqrt(((_projections(f) - g["forces_proj"]) ** 2).mean() / g["forces_sumsq"])
    e, eo = E.total_energy(t), float(g["terms"].sum())
    print(f"mixed 999k (tau 1e-5): {its} iterations, energy {abs(e / eo - 1):.2e} relative, forces {rel_sample:.2e} "
          f"rel-RMS (sample) / {rel_proj:.2e} (projections) vs the FP64 oracle at tau 1e-10; device memory "
          f"{used:.1f} GiB; {rate:.1f} evals/s")  # fmt: skip
    assert rel_sample < 5e-5 and rel_proj < 5e-5
    assert abs(e / eo - 1) < 1e-6
