"""Shared fixtures.  `gpu` tests need an MI355X and the built HIP library; everything else runs on the CPU.

The oracle (oracle/liborc.so) is test infrastructure: it is built here on demand (make -C oracle) and only
ever used as the checker.
"""

import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "pgm-md_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


def _ensure_oracle():
    lib = os.path.join(ROOT, "oracle", "liborc.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "liborc.so"], check=True, capture_output=True)


_ensure_oracle()


@pytest.fixture(scope="session")
def golden():
    return np.load(os.path.join(GOLDEN, "water512_pgmjax.npz"))


def golden_system(g, level: str, **settings):
    """The pGM-JAX water512 box at one level (q / qp / qi / qpi) with the fixture's settings
    (tests/openmm_tests.sh:18-22, tests/ref_pgmjax.py:36-40 of the reference)."""
    from pgmhip.systems import System

    cs = g["cov_strength"] if "p" in level else np.zeros_like(g["cov_strength"])
    st = dict(
        periodic=1, cutoff=0.9, beta=3.5, grid=(36, 36, 36), order=6, polarizable=int("i" in level), tol=1e-10,
        max_iter=1000, peek=0.0, predictor=0, ke=138.935458, skin=0.15,
    )  # fmt: skip
    st.update(settings)
    return System(f"w512-{level}", g["pos"], g["H"], g["q"], g["radius"], g["alpha"], g["cov_atom"], g["cov_partner"],
                  cs, st)  # fmt: skip


def triclinic_waters(nmol: int, seed: int, periodic: bool = True):
    """Random water-like molecules in the triclinic box of the reference's finite-difference test
    (core/tests/test_core.cpp:12-38,43-56): O-H 0.096 nm in random directions, q -0.8/0.4, R 0.0605/0.0536,
    alpha 1.1e-3/0.33e-3, covalent c(O->H) = -0.019, c(H->O) = 0.0086; cutoff 0.85, beta 5, 48^3, order 6."""
    from pgmhip.systems import System

    box = np.array([[2.0, 0, 0], [-0.5, 1.9, 0], [-0.4, -0.7, 1.8]])
    rng = np.random.default_rng(seed)
    pos, q, R, al, ca, cp, cs = [], [], [], [], [], [], []
    for m in range(nmol):
        f = rng.random(3)
        o = f @ box
        base = len(pos)
        pos.append(o)
        for h in range(2):
            th, ph = 2 * np.pi * rng.random(), np.arccos(2 * rng.random() - 1)
            pos.append(o + 0.096 * np.array([np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)]))
        q += [-0.8, 0.4, 0.4]
        R += [0.0605, 0.0536, 0.0536]
        al += [1.1e-3, 0.33e-3, 0.33e-3]
        for h in (1, 2):
            ca += [base, base + h]
            cp += [base + h, base]
            cs += [-0.019, 0.0086]
    st = dict(
        periodic=int(periodic), cutoff=0.85, beta=5.0, grid=(48, 48, 48), order=6, polarizable=1, tol=1e-13,
        max_iter=2000, peek=0.0, predictor=0, ke=138.935456, skin=0.1,
    )  # fmt: skip
    return System(f"tric{nmol}", np.asarray(pos), box, np.asarray(q), np.asarray(R), np.asarray(al),
                  np.asarray(ca, np.int32), np.asarray(cp, np.int32), np.asarray(cs), st)  # fmt: skip


def rel_rms(f, fref):
    f, fref = np.asarray(f), np.asarray(fref)
    return float(np.sqrt(((f - fref) ** 2).mean()) / np.sqrt((fref**2).mean()))
