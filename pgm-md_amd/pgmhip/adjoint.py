"""Per-type parameter gradients of pGM energies and forces as a torch.autograd.Function.

`PgmModel` binds one engine to one system's topology and its parameter types (`prmtop.parameter_types`);
calling it with per-type parameter tensors and positions returns the energy and forces, and differentiating any
loss of them with respect to the per-type charges, radii, polarizabilities and covalent strengths runs the
engine's hand-written adjoint (pgm_backward_device): a second CG solve with the symmetric induction matrix of
the forward solve and a recomputation of the PME grids.  Positions are data here: no gradient flows to them.

The reference has no parameter derivatives; its parameters come from the prmtop sections that
openmm/python/pgm_prmtop.py:54-115 reads (`prmtop.pgm_parameters`).
"""

from __future__ import annotations

import numpy as np
import torch

from .engine import Engine, PgmError


class PgmModel:
    """One system: an engine, the atom / covalent type maps and the box.

    engine: an `Engine` (any precision) whose settings fit the box; atom_type (n,), cov_atom, cov_partner,
    cov_type (ncov,): the topology and the type of every covalent entry; box (3, 3) rows.
    """

    def __init__(self, engine: Engine, atom_type, cov_atom, cov_partner, cov_type, box):
        self.engine = engine
        self.atom_type = np.ascontiguousarray(atom_type, np.int64)
        self.cov_atom = np.ascontiguousarray(cov_atom, np.int32)
        self.cov_partner = np.ascontiguousarray(cov_partner, np.int32)
        self.cov_type = np.ascontiguousarray(cov_type, np.int64)
        self.box = np.ascontiguousarray(box, np.float64).reshape(3, 3)
        self.n = len(self.atom_type)
        self.ncov = len(self.cov_type)
        self._bound = None
        self._index = {}

    @classmethod
    def from_types(cls, engine: Engine, types: dict, box) -> "PgmModel":
        return cls(engine, types["atom_type"], types["cov_atom"], types["cov_partner"], types["cov_type"], box)

    def index(self, device) -> tuple:
        key = str(device)
        if key not in self._index:
            self._index[key] = (
                torch.as_tensor(self.atom_type, device=device),
                torch.as_tensor(self.cov_type, device=device),
            )
        return self._index[key]

    def bind(self, q_t, radius_t, alpha_t, cov_t):
        """Expand per-type values to the engine's per-atom parameters (host copies; skipped if unchanged)."""
        vals = tuple(np.asarray(t.detach().to("cpu", torch.float64)) for t in (q_t, radius_t, alpha_t, cov_t))
        if self._bound is not None and all(np.array_equal(a, b) for a, b in zip(vals, self._bound)):
            return
        tq, tr, ta, tc = vals
        self.engine.set_atoms(tq[self.atom_type], tr[self.atom_type], ta[self.atom_type])
        self.engine.set_covalent(self.cov_atom, self.cov_partner, tc[self.cov_type] if self.ncov else np.zeros(0))
        self._bound = vals

    def __call__(self, q_t, radius_t, alpha_t, cov_t, pos):
        """(energy [kJ/mol] 0-d tensor, forces [kJ/mol/nm] (n, 3)) on pos.device, FP64."""
        return PgmFunction.apply(q_t, radius_t, alpha_t, cov_t, pos, self)


class PgmFunction(torch.autograd.Function):
    """Energy and forces of a `PgmModel` as functions of the per-type parameters (see the module docstring)."""

    @staticmethod
    def forward(ctx, q_t, radius_t, alpha_t, cov_t, pos, model: PgmModel):
        if not pos.is_cuda:
            raise PgmError("PgmFunction needs positions on the GPU (no CPU path exists)")
        if pos.shape != (model.n, 3):
            raise PgmError(f"positions of shape {tuple(pos.shape)}, model has {model.n} atoms")
        model.bind(q_t, radius_t, alpha_t, cov_t)
        pos64 = pos.detach().to(torch.float64).contiguous()
        forces = torch.zeros_like(pos64)
        stream = torch.cuda.current_stream(pos.device).cuda_stream
        # same-step evaluation: repeated calls at one geometry start from the last solution
        terms = model.engine.compute_device_f64(pos64.data_ptr(), model.box, False, forces.data_ptr(), stream)
        energy = torch.tensor(sum(terms.values()), dtype=torch.float64, device=pos.device)
        ctx.set_materialize_grads(False)
        ctx.model = model
        ctx.params = (q_t.detach().clone(), radius_t.detach().clone(), alpha_t.detach().clone(), cov_t.detach().clone())
        ctx.save_for_backward(pos64)
        return energy, forces

    @staticmethod
    def backward(ctx, g_energy, g_forces):
        model: PgmModel = ctx.model
        (pos64,) = ctx.saved_tensors
        q_t, radius_t, alpha_t, cov_t = ctx.params
        model.bind(q_t, radius_t, alpha_t, cov_t)
        dev = pos64.device
        n, ncov = model.n, model.ncov
        gq = torch.empty(n, dtype=torch.float64, device=dev)
        gr = torch.empty(n, dtype=torch.float64, device=dev)
        ga = torch.empty(n, dtype=torch.float64, device=dev)
        gc = torch.empty(max(ncov, 1), dtype=torch.float64, device=dev)
        wE = 0.0 if g_energy is None else float(g_energy)
        gF = None if g_forces is None else g_forces.to(torch.float64).contiguous()
        stream = torch.cuda.current_stream(dev).cuda_stream
        model.engine.backward_device(
            pos64.data_ptr(), model.box, None if gF is None else gF.data_ptr(), wE, gq.data_ptr(), gr.data_ptr(),
            ga.data_ptr(), gc.data_ptr(), stream,
        )  # fmt: skip
        at, ct = model.index(dev)

        def reduce(per_atom, like, index):
            return torch.zeros(like.shape, dtype=torch.float64, device=dev).index_add_(0, index, per_atom).to(like)

        grads = (
            reduce(gq, q_t, at) if ctx.needs_input_grad[0] else None,
            reduce(gr, radius_t, at) if ctx.needs_input_grad[1] else None,
            reduce(ga, alpha_t, at) if ctx.needs_input_grad[2] else None,
            reduce(gc[:ncov], cov_t, ct) if ctx.needs_input_grad[3] else None,
        )
        return (*grads, None, None)


def force_matching_loss(models, params, positions, reference_forces) -> torch.Tensor:
    """Mean over all atoms of all snapshots of |F(theta) - F_ref|^2 (kJ^2 mol^-2 nm^-2)."""
    total = 0.0
    count = 0
    for model, pos, fref in zip(models, positions, reference_forces):
        _, f = model(*params, pos)
        total = total + ((f - fref) ** 2).sum()
        count += model.n
    return total / count
