"""Seeded synthetic pGM systems: the benchmark workloads (water boxes, the planar-ring molecular liquid, the
uneven snapshot batch).

Units are the engine's (nm, e, e nm, nm^3); the Angstrom figures of the workload definitions are converted
here (1 A = 0.1 nm, A^3 -> 1e-3 nm^3, e A -> 0.1 e nm, 1/A -> 10/nm).  Every generator is deterministic in
its seed.

Choices the workload text leaves open, fixed here:
  * water: O-H 0.09572 nm, H-O-H 104.52 deg; uniform random orientations (random unit quaternions); lattice
    sites jittered uniformly by +-`jitter` nm per axis; each O carries one covalent dipole per O-H bond,
    c = +0.01 e nm toward the H (p_O += c unit(r_H - r_O)); the H atoms carry none;
  * planar-ring liquid: a regular hexagon of side 0.14 nm plus 13 substituents, each bonded to a random
    earlier atom at U(0.10, 0.15) nm in the plane; one covalent dipole per bond on the first atom toward the
    second, strength U(0, 0.015) e nm; the 450 copies share the template's parameters; Ewald beta 3.5/nm;
  * the Coulomb constant is 138.935456 kJ mol^-1 nm e^-2 (OpenMM's ONE_4PI_EPS0, pgmcore's default).
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np

KE_DEFAULT = 138.935456
R_OH = 0.09572
HOH_DEG = 104.52


@dataclasses.dataclass
class System:
    """Positions, box (rows a, b, c), per-atom parameters, covalent dipoles and engine settings."""

    name: str
    pos: np.ndarray  # (n, 3) nm
    box: np.ndarray  # (3, 3) rows
    q: np.ndarray
    radius: np.ndarray
    alpha: np.ndarray
    cov_atom: np.ndarray
    cov_partner: np.ndarray
    cov_strength: np.ndarray
    settings: dict

    @property
    def n(self) -> int:
        return len(self.q)

    def subset_settings(self, **kw) -> dict:
        s = dict(self.settings)
        s.update(kw)
        return s


def random_rotations(rng: np.random.Generator, m: int) -> np.ndarray:
    """m uniformly distributed rotation matrices (Shoemake's random unit quaternions)."""
    u1, u2, u3 = rng.random(m), rng.random(m), rng.random(m)
    a = np.sqrt(1 - u1)
    b = np.sqrt(u1)
    x, y = a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2)
    z, w = b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)
    R = np.empty((m, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def water_template() -> np.ndarray:
    """O at the origin, the two H in the xz plane."""
    h = math.radians(HOH_DEG) / 2
    return np.array(
        [[0.0, 0.0, 0.0], [R_OH * math.sin(h), 0.0, R_OH * math.cos(h)], [-R_OH * math.sin(h), 0.0, R_OH * math.cos(h)]]
    )


def water_box(
    nmol: int, L: float, nside: int, seed: int, jitter: float, name: str, settings: dict,
    q=(-0.8, 0.4), alpha=(1.0e-3, 0.4e-3), radius=(0.075, 0.055), cov=0.01,
) -> System:  # fmt: skip
    """nmol rigid waters on nmol seeded sites of an nside^3 lattice of a cubic box of edge L [nm]."""
    rng = np.random.default_rng(seed)
    sites = np.stack(np.meshgrid(*[np.arange(nside)] * 3, indexing="ij"), -1).reshape(-1, 3)
    if nmol > len(sites):
        raise ValueError("more molecules than lattice sites")
    keep = np.sort(rng.choice(len(sites), nmol, replace=False)) if nmol < len(sites) else np.arange(nmol)
    centers = (sites[keep] + 0.5) * (L / nside) + rng.uniform(-jitter, jitter, (nmol, 3))
    R = random_rotations(rng, nmol)
    tmpl = water_template()
    pos = centers[:, None, :] + np.einsum("mij,aj->mai", R, tmpl)
    pos = pos.reshape(-1, 3)
    n = 3 * nmol
    qv = np.tile([q[0], q[1], q[1]], nmol).astype(float)
    av = np.tile([alpha[0], alpha[1], alpha[1]], nmol).astype(float)
    rv = np.tile([radius[0], radius[1], radius[1]], nmol).astype(float)
    o = 3 * np.arange(nmol)
    cov_atom = np.repeat(o, 2).astype(np.int32)
    cov_partner = np.stack([o + 1, o + 2], 1).reshape(-1).astype(np.int32)
    cov_strength = np.full(2 * nmol, cov, float)
    assert len(qv) == n
    return System(name, pos, np.diag([L, L, L]).astype(float), qv, rv, av, cov_atom, cov_partner, cov_strength, settings)


def base_settings(**kw) -> dict:
    s = dict(
        periodic=1, cutoff=0.9, beta=3.5, grid=(32, 32, 32), order=6, polarizable=1, tol=1e-5, max_iter=100,
        peek=0.65, predictor=2, ke=KE_DEFAULT, skin=0.15,
    )  # fmt: skip
    s.update(kw)
    return s


def config0(seed: int = 0) -> System:
    """512 waters (1536 atoms), 2.48 nm cubic box, 8^3 jittered lattice; 32^3 PME, tau 1e-6."""
    return water_box(512, 2.48, 8, seed, 0.03, "water512", base_settings(grid=(32, 32, 32), tol=1e-6))


def ring_template(rng: np.random.Generator):
    """A rigid planar 19-atom molecule: a hexagon (side 0.14 nm) and 13 in-plane substituents."""
    ang = np.arange(6) * np.pi / 3
    atoms = [np.array([0.14 * np.cos(a), 0.14 * np.sin(a), 0.0]) for a in ang]
    bonds = [(k, (k + 1) % 6) for k in range(6)]
    while len(atoms) < 19:
        parent = int(rng.integers(len(atoms)))
        length = rng.uniform(0.10, 0.15)
        phi = rng.uniform(0, 2 * np.pi)
        cand = atoms[parent] + length * np.array([np.cos(phi), np.sin(phi), 0.0])
        d = np.linalg.norm(np.asarray(atoms) - cand, axis=1)
        d[parent] = 1.0
        if d.min() >= 0.11:
            atoms.append(cand)
            bonds.append((parent, len(atoms) - 1))
    xyz = np.asarray(atoms)
    return xyz - xyz.mean(0), bonds


def config1(seed: int = 1) -> System:
    """450 copies of a seeded planar 19-atom molecule (8550 atoms) in a 4.5 nm box; 48^3 PME, tau 1e-5."""
    rng = np.random.default_rng(seed)
    tmpl, bonds = ring_template(rng)
    na = len(tmpl)
    q = rng.normal(0.0, 0.3, na)
    q -= q.mean()
    alpha = rng.uniform(0.3e-3, 1.8e-3, na)
    radius = rng.uniform(0.05, 0.09, na)
    strength = rng.uniform(0.0, 0.015, len(bonds))
    L, nside, nmol = 4.5, 8, 450
    sites = np.stack(np.meshgrid(*[np.arange(nside)] * 3, indexing="ij"), -1).reshape(-1, 3)
    keep = np.sort(rng.choice(len(sites), nmol, replace=False))
    centers = (sites[keep] + 0.5) * (L / nside)
    placed = np.zeros((0, 3))
    mols = []
    for m in range(nmol):
        for attempt in range(2000):
            R = random_rotations(rng, 1)[0]
            xyz = centers[m] + tmpl @ R.T
            if len(placed) == 0:
                break
            d = xyz[:, None, :] - placed[None, :, :]
            d -= L * np.round(d / L)
            if np.sqrt((d * d).sum(-1)).min() >= (0.12 if attempt < 1000 else 0.09):
                break
        else:
            raise RuntimeError("could not place molecule without overlap")
        mols.append(xyz)
        placed = np.concatenate([placed, xyz])
    pos = np.concatenate(mols)
    off = na * np.arange(nmol)
    cov_atom, cov_partner, cov_strength = [], [], []
    for m in range(nmol):
        for k, (i, j) in enumerate(bonds):
            cov_atom.append(off[m] + i)
            cov_partner.append(off[m] + j)
            cov_strength.append(strength[k])
    st = base_settings(grid=(48, 48, 48), tol=1e-5)
    return System(
        "ring8550", pos, np.diag([L, L, L]), np.tile(q, nmol), np.tile(radius, nmol), np.tile(alpha, nmol),
        np.asarray(cov_atom, np.int32), np.asarray(cov_partner, np.int32), np.asarray(cov_strength), st,
    )  # fmt: skip


def config2(seed: int = 2) -> System:
    """32,000 waters (96,000 atoms) in a 6.84 nm box (3x liquid density): 32^3 lattice at 0.21375 nm, 32,000
    seeded sites kept, jitter +-0.02 nm; 72^3 PME, tau 1e-4."""
    return water_box(32000, 6.84, 32, seed, 0.02, "water96k", base_settings(grid=(72, 72, 72), tol=1e-4))


def config4(seed: int = 4) -> System:
    """333,000 waters (999,000 atoms) in a 21.54 nm box: 70^3 lattice, seeded 333,000 kept; 216^3 PME, tau 1e-5."""
    return water_box(333000, 21.54, 70, seed, 0.03, "water999k", base_settings(grid=(216, 216, 216), tol=1e-5))


def smooth_grid(L: float, spacing: float = 0.08) -> int:
    """The smallest 2^a 3^b >= L / spacing."""
    target = math.ceil(L / spacing)
    best = None
    for a in range(0, 12):
        for b in range(0, 8):
            v = 2**a * 3**b
            if v >= target and (best is None or v < best):
                best = v
    return best


def config3_snapshot(k: int, seed: int = 3) -> System:
    """Snapshot k of the 64-snapshot batch: n_mol ~ U{256..2048} (seeded), box at 1 g/cm^3."""
    rng = np.random.default_rng(seed)
    nmols = rng.integers(256, 2049, 64)
    nmol = int(nmols[k])
    L = (nmol / 33.429) ** (1.0 / 3.0)
    nside = math.ceil(nmol ** (1.0 / 3.0))
    K = smooth_grid(L)
    return water_box(nmol, L, nside, 1000 + k, 0.02, f"snap{k}", base_settings(grid=(K, K, K), tol=1e-8))


def perturb(sys: System, seed: int, sigma: float) -> np.ndarray:
    """Positions displaced by N(0, sigma) per component (earlier MD-like geometries for predictor tests)."""
    rng = np.random.default_rng(seed)
    return sys.pos + rng.normal(0.0, sigma, sys.pos.shape)


def rigid_moves(sys: System, seed: int, step: float, nsteps: int):
    """nsteps successive geometries, each molecule-agnostic atom moved by a smooth seeded velocity * step."""
    rng = np.random.default_rng(seed)
    v = rng.normal(0.0, 1.0, sys.pos.shape)
    v /= np.sqrt((v * v).sum(1, keepdims=True)).mean()
    return [sys.pos + (s + 1) * step * v for s in range(nsteps)]
