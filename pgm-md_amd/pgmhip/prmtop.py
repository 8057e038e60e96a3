"""AMBER prmtop reader for the pmemd-pgm POL_GAUSS_* sections, in engine units.

The file format (AMBER's "new" prmtop): directive lines start with '%'.  ``%FLAG name`` opens a section,
``%FORMAT(<count><type><width>[.<decimals>])`` gives its fixed-width field layout (type a: text, I: integer,
E/F: real), ``%COMMENT`` and ``%VERSION`` lines carry no data; every other line holds up to <count> fields of
<width> columns.  The pGM sections are the ones the reference's Python layer reads
(openmm/python/pgm_prmtop.py:54-115): per-atom monopoles [e], radii [A] and polarizabilities [A^3], the number
of covalent dipoles each atom carries, their partners (1-based) and strengths [e A], listed atom after atom.
Engine units: nm, nm^3, e nm.  The OpenMM System builder of that file is not part of this engine.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# the pGM sections and the factor from the file's units to the engine's (A -> nm, A^3 -> nm^3 as the literal
# 1e-3, which 0.1**3 is not in binary)
_PGM_SECTIONS = {
    "POL_GAUSS_MONOPOLES_LIST": 1.0,
    "POL_GAUSS_RADII_LIST": 0.1,
    "POL_GAUSS_POLARIZABILITY_LIST": 1e-3,
    "POL_GAUSS_COVALENT_POINTERS_LIST": None,
    "POL_GAUSS_COVALENT_ATOMS_LIST": None,
    "POL_GAUSS_COVALENT_DIPOLES_LIST": 0.1,
}


@dataclass(frozen=True)
class FieldLayout:
    """One %FORMAT specification: fields per line, field type ('a', 'i' or 'f') and field width."""

    per_line: int
    kind: str
    width: int

    @staticmethod
    def parse(spec: str) -> "FieldLayout":
        """Parses the text after %FORMAT, e.g. '(10I8)', '(5E16.8)', '(20a4)', '(i5)' (no count: one field)."""
        body = spec.strip()
        if not (body.startswith("(") and ")" in body):
            raise ValueError(f"prmtop: bad %FORMAT specification {spec!r}")
        body = body[1 : body.index(")")].strip()
        k = 0
        while k < len(body) and body[k].isdigit():
            k += 1
        if k == len(body):
            raise ValueError(f"prmtop: bad %FORMAT specification {spec!r}")
        letter = body[k].lower()
        kinds = {"a": "a", "i": "i", "e": "f", "f": "f", "d": "f", "g": "f"}
        if letter not in kinds:
            raise ValueError(f"prmtop: unknown field type {body[k]!r} in {spec!r}")
        width_text = body[k + 1 :].split(".")[0]
        if not width_text.isdigit() or int(width_text) <= 0:
            raise ValueError(f"prmtop: bad field width in {spec!r}")
        return FieldLayout(int(body[:k]) if k else 1, kinds[letter], int(width_text))

    def fields(self, line: str) -> list[str]:
        """The non-blank fields of one data line, in order."""
        out = []
        for col in range(0, min(len(line), self.per_line * self.width), self.width):
            piece = line[col : col + self.width].strip()
            if piece:
                out.append(piece)
        return out


class Prmtop:
    """A prmtop file indexed by section; values are converted when asked for.

    ``Prmtop(path).get(name)`` returns a numpy int64 or float64 array for numeric sections and a list of str
    for text sections; ``name in p`` tests presence; ``p.names`` lists the sections in file order.
    """

    def __init__(self, path: str):
        with open(path) as fh:
            self._lines = fh.read().splitlines()
        self._path = path
        self._index: dict[str, tuple[str | None, int, int]] = {}
        self._scan()

    def _scan(self) -> None:
        # a section spans from its %FLAG line to the next one; its data lines are those not starting with '%'
        flags = [k for k, ln in enumerate(self._lines) if ln.startswith("%FLAG")]
        for pos, first in enumerate(flags):
            last = flags[pos + 1] if pos + 1 < len(flags) else len(self._lines)
            head = self._lines[first].split()
            if len(head) < 2:
                raise ValueError(f"{self._path}:{first + 1}: %FLAG without a section name")
            spec = next((ln[len("%FORMAT") :] for ln in self._lines[first + 1 : last] if ln.startswith("%FORMAT")),
                        None)  # fmt: skip
            self._index[head[1]] = (spec, first + 1, last)

    @property
    def names(self) -> list[str]:
        return list(self._index)

    def __contains__(self, name: str) -> bool:
        return name in self._index

    def get(self, name: str):
        spec, first, last = self._index[name]
        data = [ln for ln in self._lines[first:last] if not ln.startswith("%")]
        if spec is None:  # no %FORMAT: the raw lines
            return data
        layout = FieldLayout.parse(spec)
        tokens = [t for ln in data for t in layout.fields(ln)]
        if layout.kind == "a":
            return tokens
        if layout.kind == "i":
            return np.array([int(t) for t in tokens], dtype=np.int64)
        return np.array([float(t) for t in tokens], dtype=np.float64)


def read_sections(path: str, names: tuple[str, ...] | None = None) -> dict:
    """Sections of a prmtop (all, or those named): numpy arrays for numbers, lists of str for text."""
    p = Prmtop(path)
    wanted = p.names if names is None else [k for k in names if k in p]
    return {k: p.get(k) for k in wanted}


def pgm_parameters(path: str) -> dict:
    """q [e], radius [nm], alpha [nm^3], cov_atom / cov_partner (0-based), cov_strength [e nm]."""
    p = Prmtop(path)
    missing = [k for k in _PGM_SECTIONS if k not in p]
    if missing:
        raise ValueError(f"{path}: pGM sections missing: {', '.join(missing)}")
    v = {k: p.get(k) for k in _PGM_SECTIONS}
    n = len(v["POL_GAUSS_MONOPOLES_LIST"])
    per_atom = v["POL_GAUSS_COVALENT_POINTERS_LIST"]
    partner = v["POL_GAUSS_COVALENT_ATOMS_LIST"] - 1
    strength = v["POL_GAUSS_COVALENT_DIPOLES_LIST"]
    for k in ("POL_GAUSS_RADII_LIST", "POL_GAUSS_POLARIZABILITY_LIST", "POL_GAUSS_COVALENT_POINTERS_LIST"):
        if len(v[k]) != n:
            raise ValueError(f"{path}: {k} has {len(v[k])} entries for {n} atoms")
    ncov = int(per_atom.sum())
    if ncov != len(partner) or ncov != len(strength) or np.any(per_atom < 0):
        raise ValueError(f"{path}: the covalent counts per atom add to {ncov}, the partner list has {len(partner)}"
                         f" and the strength list {len(strength)} entries")  # fmt: skip
    if ncov and (partner.min() < 0 or partner.max() >= n):
        raise ValueError(f"{path}: a covalent partner index is outside 1..{n}")
    carrier = np.repeat(np.arange(n, dtype=np.int32), per_atom)
    return {
        "q": v["POL_GAUSS_MONOPOLES_LIST"] * _PGM_SECTIONS["POL_GAUSS_MONOPOLES_LIST"],
        "radius": v["POL_GAUSS_RADII_LIST"] * _PGM_SECTIONS["POL_GAUSS_RADII_LIST"],
        "alpha": v["POL_GAUSS_POLARIZABILITY_LIST"] * _PGM_SECTIONS["POL_GAUSS_POLARIZABILITY_LIST"],
        "cov_atom": carrier,
        "cov_partner": partner.astype(np.int32),
        "cov_strength": strength * _PGM_SECTIONS["POL_GAUSS_COVALENT_DIPOLES_LIST"],
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
