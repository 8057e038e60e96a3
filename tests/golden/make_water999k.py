"""Writes tests/golden/water999k_oracle.npz: the FP64 oracle's evaluation of the 999,000-atom water box.

The box is pgmhip.systems.config4() (333,000 waters in a 21.54 nm cube, 216^3 PME, order 6), evaluated by
oracle/pgm_oracle.cpp at tau = 1e-10 with no predictor and no peek (the oracle settings of
core/tests/run_dump.cpp:40-43).  One evaluation takes about 3.5 minutes and 17 GB on one core, so the fixture
keeps what checks a GPU evaluation of the same box without the oracle at hand:

- the five energy terms, the CG iterations and final residual, the strain S = dE/d eps;
- forces and induced dipoles of 3,000 seeded atoms (sample);
- projections of all 999,000 forces and induced dipoles on 8 seeded unit-variance Gaussian vectors w_k
  (w regenerated from seed 999): for a difference dF, E[(dF . w)^2] = |dF|^2, so the projections bound the
  error of every atom, not only of the sample;
- sum |F|^2 and sum |mu|^2 over all atoms.

    python tests/golden/make_water999k.py        (needs oracle/liborc.so: make -C oracle)
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "pgm-md_amd"))

import oracle as O  # noqa: E402
from pgmhip import systems as S  # noqa: E402

NSAMPLE, NPROJ, SAMPLE_SEED, PROJ_SEED = 3000, 8, 998, 999


def projection_vectors(n):
    return np.random.default_rng(PROJ_SEED).standard_normal((NPROJ, n, 3))


def sample_atoms(n):
    return np.sort(np.random.default_rng(SAMPLE_SEED).choice(n, NSAMPLE, replace=False))


def main():
    sy = S.config4()
    sy.settings.update(tol=1e-10, max_iter=500, peek=0.0, predictor=0)
    t0 = time.time()
    o = O.Oracle(sy)
    _, f = o.compute(sy.pos, sy.box)
    mu = o.induced()
    secs = time.time() - t0
    w = projection_vectors(sy.n)
    idx = sample_atoms(sy.n)
    t = o.terms()
    np.savez_compressed(
        os.path.join(HERE, "water999k_oracle.npz"),
        terms=np.array([t[k] for k in ("direct", "reciprocal", "self", "background", "polarization")]),
        iterations=np.int32(o.iterations()), residual=np.float64(o.residual()), strain=o.strain(),
        sample=idx.astype(np.int32), forces_sample=f[idx], induced_sample=mu[idx],
        forces_proj=np.einsum("kna,na->k", w, f), induced_proj=np.einsum("kna,na->k", w, mu),
        forces_sumsq=np.float64((f * f).sum()), induced_sumsq=np.float64((mu * mu).sum()),
        seconds=np.float64(secs), n=np.int32(sy.n),
    )  # fmt: skip
    print(f"wrote water999k_oracle.npz: {o.iterations()} iterations, {secs:.0f} s on one core")


if __name__ == "__main__":
    main()
