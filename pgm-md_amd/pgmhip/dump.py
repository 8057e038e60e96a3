"""The pGM dump of one evaluation (GROMACS `GMX_PGM_DUMP`), and its recomputation with this engine.

A dump holds everything needed to recompute one call of the electrostatics: settings, the five energy terms, CG
iterations and residual, the box, the virial Xi = S^T/2, per atom (global order) position, charge, radius,
polarizability, force, induced and permanent dipole, and the covalent dipoles; optionally the GVDW outputs. The
layout is the reference's writer, `gromacs/src/pgm/pgm.cpp:2613-2681`, and reader, `core/tests/run_dump.cpp:30-92`:
lines starting with '#' are comments, numbers are written with 17 significant digits.

    python -m pgmhip.dump DUMP [tol]

recomputes a dump with the FP64 engine (tau = tol, default 1e-10, no predictor, no peek, as
`core/tests/run_dump.cpp:40-43`) and prints the differences, as `run_dump` does.
"""

from __future__ import annotations

import dataclasses
import sys

import numpy as np

from . import systems as S


@dataclasses.dataclass
class Dump:
    beta: float
    cutoff: float
    grid: tuple
    order: int
    ke: float
    terms: dict  # direct, reciprocal, self, background, polarization [kJ/mol]
    iterations: int
    residual: float
    box: np.ndarray  # (3, 3) rows [nm]
    virial: np.ndarray  # Xi = S^T / 2 [kJ/mol]
    pos: np.ndarray  # (n, 3) [nm], global order
    q: np.ndarray
    radius: np.ndarray
    alpha: np.ndarray
    forces: np.ndarray  # (n, 3) [kJ/mol/nm]
    induced: np.ndarray  # (n, 3) [e nm]
    permanent: np.ndarray  # (n, 3) [e nm]
    cov_atom: np.ndarray
    cov_partner: np.ndarray
    cov_strength: np.ndarray
    gvdw: dict | None = None  # energies (pairs, overrides, tail), virial (3, 3), forces (n, 3)

    @property
    def n(self) -> int:
        return len(self.q)

    def system(self, **settings) -> S.System:
        """The dump as a system for the engine and the oracle (periodic; tau 1e-10, no predictor, no peek)."""
        st = S.base_settings(
            cutoff=self.cutoff, beta=self.beta, grid=tuple(self.grid), order=self.order, ke=self.ke, tol=1e-10,
            max_iter=500, peek=0.0, predictor=0, polarizable=int(bool((self.alpha > 0).any())),
        )  # fmt: skip
        st.update(settings)
        return S.System("dump", self.pos.copy(), self.box.copy(), self.q.copy(), self.radius.copy(),
                        self.alpha.copy(), self.cov_atom.astype(np.int32), self.cov_partner.astype(np.int32),
                        self.cov_strength.copy(), st)  # fmt: skip


TERMS = ("direct", "reciprocal", "self", "background", "polarization")


def _data_lines(path):
    with open(path) as fh:
        for line in fh:
            if line.strip() and not line.startswith("#"):
                yield line.split()


def read(path: str) -> Dump:
    """Read a dump (the layout of `core/tests/run_dump.cpp:30-92`, plus the optional GVDW block)."""
    it = _data_lines(path)
    h = next(it)
    n, ncov = int(h[0]), int(h[1])
    beta, rc = float(h[2]), float(h[3])
    grid = (int(h[4]), int(h[5]), int(h[6]))
    order, ke = int(h[7]), float(h[8])
    t = next(it)
    terms = {k: float(v) for k, v in zip(TERMS, t[:5])}
    iterations, residual = int(t[5]), float(t[6])
    box = np.array([[float(v) for v in next(it)[:3]] for _ in range(3)])
    vir = np.array([[float(v) for v in next(it)[:3]] for _ in range(3)])
    atoms = np.array([[float(v) for v in next(it)[:15]] for _ in range(n)]).reshape(n, 15)
    cov = [next(it)[:3] for _ in range(ncov)]
    gvdw = None
    rest = list(it)
    if rest:
        e = [float(v) for v in rest[0][:3]]
        gv = np.array([[float(v) for v in row[:3]] for row in rest[1:4]])
        gf = np.array([[float(v) for v in row[:3]] for row in rest[4 : 4 + n]])
        gvdw = {"energy_pairs": e[0], "energy_overrides": e[1], "energy_tail": e[2], "virial": gv, "forces": gf}
    return Dump(
        beta=beta, cutoff=rc, grid=grid, order=order, ke=ke, terms=terms, iterations=iterations, residual=residual,
        box=box, virial=vir, pos=atoms[:, 0:3].copy(), q=atoms[:, 3].copy(), radius=atoms[:, 4].copy(),
        alpha=atoms[:, 5].copy(), forces=atoms[:, 6:9].copy(), induced=atoms[:, 9:12].copy(),
        permanent=atoms[:, 12:15].copy(), cov_atom=np.array([int(c[0]) for c in cov], np.int64),
        cov_partner=np.array([int(c[1]) for c in cov], np.int64), cov_strength=np.array([float(c[2]) for c in cov]),
        gvdw=gvdw,
    )  # fmt: skip


def write(path: str, d: Dump) -> None:
    """Write a dump in the writer's layout (`gromacs/src/pgm/pgm.cpp:2613-2681`)."""
    g = "%.17g"
    with open(path, "w") as fh:
        fh.write("# pGM dump: n ncov beta rc K1 K2 K3 order ke\n")
        fh.write(f"{d.n} {len(d.cov_atom)} {d.beta:.17g} {d.cutoff:.17g} {d.grid[0]} {d.grid[1]} {d.grid[2]} "
                 f"{d.order} {d.ke:.17g}\n")  # fmt: skip
        fh.write("# terms [kJ/mol]: direct reciprocal self background polarization; iterations residual\n")
        fh.write(" ".join(g % d.terms[k] for k in TERMS) + f" {d.iterations} {d.residual:.6g}\n")
        fh.write("# box rows [nm]\n")
        for a in range(3):
            fh.write(" ".join(g % v for v in d.box[a]) + "\n")
        fh.write("# virial (Xi = S^T/2) [kJ/mol]\n")
        for a in range(3):
            fh.write(" ".join(g % v for v in d.virial[a]) + "\n")
        fh.write("# atom (global order): x(3) [nm] q radius alpha, pGM force (3) [kJ/mol/nm], induced dipole (3) "
                 "[e nm], permanent dipole (3)\n")  # fmt: skip
        for i in range(d.n):
            row = [*d.pos[i], d.q[i], d.radius[i], d.alpha[i], *d.forces[i], *d.induced[i], *d.permanent[i]]
            fh.write(" ".join(g % v for v in row) + "\n")
        fh.write("# covalent dipoles: atom partner strength [e nm] (0-based, global)\n")
        for a, b, c in zip(d.cov_atom, d.cov_partner, d.cov_strength):
            fh.write(f"{int(a)} {int(b)} {c:.17g}\n")
        if d.gvdw is not None:
            gv = d.gvdw
            fh.write("# GVDW: energy (pairs) energy (overrides) energy (tail) [kJ/mol]; virial (3x3); per atom "
                     "(global order) force [kJ/mol/nm]\n")  # fmt: skip
            fh.write(" ".join(g % gv[k] for k in ("energy_pairs", "energy_overrides", "energy_tail")) + "\n")
            for a in range(3):
                fh.write(" ".join(g % v for v in gv["virial"][a]) + "\n")
            for i in range(d.n):
                fh.write(" ".join(g % v for v in gv["forces"][i]) + "\n")


def from_evaluation(system: S.System, terms: dict, forces, strain, induced, permanent, iterations: int,
                    residual: float) -> Dump:  # fmt: skip
    """A dump of one evaluation (terms in kJ/mol, forces in kJ/mol/nm, strain S_ab = dE/d eps_ab)."""
    st = system.settings
    S3 = np.asarray(strain).reshape(3, 3)
    return Dump(
        beta=st["beta"], cutoff=st["cutoff"], grid=tuple(st["grid"]), order=st["order"], ke=st["ke"],
        terms={k: terms[k] for k in TERMS}, iterations=int(iterations), residual=float(residual),
        box=np.asarray(system.box, float), virial=0.5 * S3.T, pos=np.asarray(system.pos, float),
        q=system.q, radius=system.radius, alpha=system.alpha, forces=np.asarray(forces, float),
        induced=np.asarray(induced, float), permanent=np.asarray(permanent, float), cov_atom=system.cov_atom,
        cov_partner=system.cov_partner, cov_strength=system.cov_strength,
    )  # fmt: skip


def recompute(d: Dump, tol: float = 1e-10, precision: str = "double") -> dict:
    """Evaluate the dump with this engine and compare (the quantities `core/tests/run_dump.cpp:93-128` prints)."""
    from . import engine as E

    sy = d.system(tol=tol)
    eng = E.Engine.for_system(sy, precision=precision)
    terms, f, strain = eng.compute(sy.pos, sy.box, strain=True)
    xi = 0.5 * np.asarray(strain).T
    total_d, total = sum(d.terms.values()), sum(terms[k] for k in TERMS)
    return {
        "terms": {k: (d.terms[k], terms[k]) for k in TERMS},
        "total_rel": (total_d - total) / abs(total),
        "force_max_diff": float(np.abs(d.forces - f).max()),
        "force_rms": float(np.sqrt((f * f).mean())),
        "induced_max_diff": float(np.abs(d.induced - eng.induced_dipoles()).max()),
        "permanent_max_diff": float(np.abs(d.permanent - eng.permanent_dipoles()).max()),
        "iterations": (d.iterations, eng.iterations()),
        "virial_max_diff": float(np.abs(d.virial - xi).max()),
        "virial_max": float(np.abs(xi).max()),
    }


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print("usage: python -m pgmhip.dump DUMP [tol]", file=sys.stderr)
        return 1
    r = recompute(read(argv[0]), tol=float(argv[1]) if len(argv) > 1 else 1e-10)
    for k, (a, b) in r["terms"].items():
        print(f"{k:<13s} dump {a:22.12f} engine {b:22.12f} diff {a - b:10.3e}")
    print(f"total relative difference {r['total_rel']:.3e}")
    print(f"forces max|diff| {r['force_max_diff']:.3e} (rms force {r['force_rms']:.3e});  mu max|diff| "
          f"{r['induced_max_diff']:.3e};  perm max|diff| {r['permanent_max_diff']:.3e}")  # fmt: skip
    print(f"CG iterations dump {r['iterations'][0]} engine {r['iterations'][1]}")
    print(f"virial max|diff| {r['virial_max_diff']:.3e} (max|Xi| {r['virial_max']:.3e})")
    return 0


if __name__ == "__main__":
    sys.exit(main())
