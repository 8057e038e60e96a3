"""pmemd-pgm prmtop reader: the POL_GAUSS_* sections in engine units.

Behaviour follows the reference's Python layer openmm/python/pgm_prmtop.py:54-115 (`read_sections`,
`pgm_parameters`): %FLAG sections parsed with their %FORMAT widths; Angstrom -> nm, A^3 -> nm^3,
e A -> e nm; covalent dipoles listed atom after atom (POL_GAUSS_COVALENT_POINTERS_LIST counts, partners
1-based in POL_GAUSS_COVALENT_ATOMS_LIST).  The OpenMM System builder of that file is not part of this
engine.
"""

from __future__ import annotations

import re

import numpy as np

_FMT = re.compile(r"\((\d+)([aAiIeEfF])(\d+)")


def read_sections(path: str, names: tuple[str, ...] | None = None) -> dict:
    """The %FLAG sections of a prmtop: numpy arrays for numbers, lists of str for text."""
    out: dict = {}
    name, fmt, lines = None, None, []

    def flush():
        if name is None or (names is not None and name not in names):
            return
        m = _FMT.match(fmt or "")
        if m is None:
            out[name] = list(lines)
            return
        kind, width = m.group(2).lower(), int(m.group(3))
        toks = [ln[k : k + width].strip() for ln in lines for k in range(0, len(ln), width)]
        toks = [t for t in toks if t]
        if kind == "a":
            out[name] = toks
        elif kind == "i":
            out[name] = np.array([int(t) for t in toks], dtype=np.int64)
        else:
            out[name] = np.array([float(t) for t in toks], dtype=np.float64)

    with open(path) as fh:
        for raw in fh:
            line = raw.rstrip("\n")
            if line.startswith("%FLAG"):
                flush()
                name, fmt, lines = line.split()[1], None, []
            elif line.startswith("%FORMAT"):
                fmt = line[line.index("(") :]
            elif line.startswith("%COMMENT") or line.startswith("%VERSION"):
                continue
            elif name is not None:
                lines.append(line)
        flush()
    return out


def pgm_parameters(path: str) -> dict:
    """q [e], radius [nm], alpha [nm^3], cov_atom / cov_partner (0-based), cov_strength [e nm]."""
    s = read_sections(path)
    if "POL_GAUSS_MONOPOLES_LIST" not in s:
        raise ValueError(f"{path}: not a pGM prmtop (no POL_GAUSS_MONOPOLES_LIST)")
    counts = s["POL_GAUSS_COVALENT_POINTERS_LIST"].astype(np.int64)
    partners = s["POL_GAUSS_COVALENT_ATOMS_LIST"].astype(np.int64) - 1
    if counts.sum() != len(partners):
        raise ValueError(f"{path}: covalent pointer counts do not match the partner list")
    return {
        "q": np.asarray(s["POL_GAUSS_MONOPOLES_LIST"], float),
        "radius": np.asarray(s["POL_GAUSS_RADII_LIST"], float) * 0.1,
        "alpha": np.asarray(s["POL_GAUSS_POLARIZABILITY_LIST"], float) * 1e-3,
        "cov_atom": np.repeat(np.arange(len(counts)), counts).astype(np.int32),
        "cov_partner": partners.astype(np.int32),
        "cov_strength": np.asarray(s["POL_GAUSS_COVALENT_DIPOLES_LIST"], float) * 0.1,
    }


def parameter_types(params: dict, atom_names=None) -> dict:
    """Group the per-atom parameters of `pgm_parameters` into types for fitting.

    Atoms that share (name, q, radius, alpha) share an atom type (the name only when `atom_names` is given,
    e.g. the prmtop's AMBER_ATOM_TYPE); covalent entries that share (carrier type, partner type, strength)
    share a covalent type.  Returns the type index arrays ("atom_type" (n,), "cov_type" (ncov,)) and the
    per-type values ("q", "radius", "alpha" per atom type, "cov_strength" per covalent type); expanding the
    per-type values by the index arrays gives back the per-atom parameters exactly.
    """
    q = np.asarray(params["q"], float)
    r = np.asarray(params["radius"], float)
    a = np.asarray(params["alpha"], float)
    names = [""] * len(q) if atom_names is None else [str(x) for x in atom_names]
    if len(names) != len(q):
        raise ValueError("one atom name per atom")
    keys: dict = {}
    atom_type = np.empty(len(q), np.int64)
    for i in range(len(q)):
        atom_type[i] = keys.setdefault((names[i], q[i], r[i], a[i]), len(keys))
    tq, tr, ta = (np.zeros(len(keys)) for _ in range(3))
    for (_, qi, ri, ai), t in keys.items():
        tq[t], tr[t], ta[t] = qi, ri, ai
    ca = np.asarray(params["cov_atom"], np.int64)
    cp = np.asarray(params["cov_partner"], np.int64)
    cs = np.asarray(params["cov_strength"], float)
    ckeys: dict = {}
    cov_type = np.empty(len(ca), np.int64)
    for m in range(len(ca)):
        cov_type[m] = ckeys.setdefault((int(atom_type[ca[m]]), int(atom_type[cp[m]]), cs[m]), len(ckeys))
    tc = np.zeros(len(ckeys))
    for (_, _, sm), t in ckeys.items():
        tc[t] = sm
    return {
        "atom_type": atom_type, "cov_type": cov_type, "q": tq, "radius": tr, "alpha": ta, "cov_strength": tc,
        "cov_atom": ca.astype(np.int32), "cov_partner": cp.astype(np.int32),
    }  # fmt: skip
