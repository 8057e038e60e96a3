"""Python binding of the HIP engine (ctypes over the C ABI of include/pgm_hip.h).

`Engine` mirrors pgmcuda::Engine (core/cuda/pgm_cuda.h:47-127): set_atoms, set_covalent, compute,
compute_f32 + forces, compute_device (float4 positions, fixed-point forces), iterations, residual,
induced/permanent dipoles, reset_history, list_builds, num_candidates.  Errors raise `PgmError` with the
engine's message.  The library must be built (pgm-md_amd/build_hip.py); a missing library is an error, never a
fallback.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libpgm_hip.so")

PGM_DOUBLE, PGM_MIXED = 0, 1


class PgmError(RuntimeError):
    pass


class Settings(C.Structure):
    """pgm_settings (include/pgm_hip.h)."""

    _fields_ = [
        ("cutoff", C.c_double),
        ("beta", C.c_double),
        ("grid", C.c_int * 3),
        ("order", C.c_int),
        ("tol", C.c_double),
        ("max_iter", C.c_int),
        ("peek", C.c_double),
        ("predictor", C.c_int),
        ("ke", C.c_double),
        ("device", C.c_int),
        ("skin", C.c_double),
        ("precision", C.c_int),
        ("polarizable", C.c_int),
    ]


class Terms(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("direct", "reciprocal", "self", "background", "polarization", "gvdw")]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}

    @property
    def total(self) -> float:
        return sum(getattr(self, k) for k, _ in self._fields_)


_lib = None
EXPORTS = [
    "pgm_default_settings", "pgm_create", "pgm_destroy", "pgm_last_error", "pgm_set_atoms", "pgm_set_covalent",
    "pgm_compute", "pgm_compute_f32", "pgm_forces", "pgm_compute_device", "pgm_launch_device", "pgm_finish_device",
    "pgm_compute_device_f64", "pgm_iterations", "pgm_residual", "pgm_induced_dipoles", "pgm_permanent_dipoles",
    "pgm_reset_history", "pgm_list_builds", "pgm_num_candidates", "pgm_late_list_checks", "pgm_set_history",
    "pgm_freeze_history", "pgm_set_timing", "pgm_set_timing_phases", "pgm_get_timing", "pgm_num_pairs", "pgm_backward",
    "pgm_backward_device", "pgm_adjoint_iterations", "pgm_adjoint_residual", "pgm_set_batch", "pgm_num_systems",
    "pgm_compute_batch", "pgm_compute_batch_device", "pgm_backward_batch", "pgm_backward_batch_device",
    "pgm_autotune", "pgm_set_lanes", "pgm_eval_erfc_f32", "pgm_eval_pair_kernels", "pgm_set_gvdw",
    "pgm_set_preconditioner", "pgm_set_candidates", "pgm_wait_device", "pgm_launch_pending", "pgm_set_field_otf",
    "pgm_field_otf", "pgm_field", "pgm_compute_device_f64_noalloc", "pgm_set_pme_overlap", "pgm_pme_overlap",
    "pgm_list_check_state",
    "pgm_pme_overlap_times", "pgm_setup_lanes", "pgm_set_setup_lanes",
]  # fmt: skip
PHASES = ["list", "prep", "pair_compact", "pair_setup", "pair_field", "pme_spread", "pme_fft", "pme_gather", "cg",
          "pair_final", "final"]  # fmt: skip


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libpgm_hip.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PgmError(f"{path} is not built: run pgm-md_amd/build_hip.py (no CPU fallback exists)")
    # One HIP runtime per process: PyTorch ships its own libamdhip64.so.7.  Loaded first, it serves this
    # library too (same soname); loaded after ROCm's copy, torch's HIP layer would bind to the wrong runtime
    # and find no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    lib.pgm_default_settings.argtypes = [C.POINTER(Settings)]
    lib.pgm_create.restype = vp
    lib.pgm_create.argtypes = [C.POINTER(Settings)]
    lib.pgm_destroy.argtypes = [vp]
    lib.pgm_last_error.restype = C.c_char_p
    lib.pgm_last_error.argtypes = [vp]
    lib.pgm_set_atoms.argtypes = [vp, C.c_int, vp, vp, vp, vp, C.c_int]
    lib.pgm_set_covalent.argtypes = [vp, C.c_int, vp, vp, vp]
    lib.pgm_compute.argtypes = [vp, vp, vp, C.c_int, vp, vp, C.POINTER(Terms)]
    lib.pgm_compute_f32.argtypes = [vp, vp, vp, C.c_int, vp, C.POINTER(Terms)]
    lib.pgm_forces.argtypes = [vp, vp]
    lib.pgm_compute_device.argtypes = [vp, vp, vp, C.c_int, vp, C.c_int, C.c_double, vp, C.POINTER(Terms)]
    lib.pgm_launch_device.argtypes = [vp, vp, vp, C.c_int, vp, C.c_int, C.c_double, vp]
    lib.pgm_finish_device.argtypes = [vp, C.POINTER(Terms), ip]
    lib.pgm_compute_device_f64.argtypes = [vp, vp, vp, C.c_int, vp, vp, vp, C.POINTER(Terms)]
    lib.pgm_compute_device_f64_noalloc.argtypes = [vp, vp, vp, vp, vp, vp, C.POINTER(Terms)]
    lib.pgm_iterations.argtypes = [vp]
    lib.pgm_residual.argtypes = [vp]
    lib.pgm_residual.restype = C.c_double
    lib.pgm_induced_dipoles.argtypes = [vp, vp]
    lib.pgm_permanent_dipoles.argtypes = [vp, vp]
    lib.pgm_reset_history.argtypes = [vp]
    lib.pgm_list_builds.argtypes = [vp]
    lib.pgm_list_builds.restype = C.c_long
    lib.pgm_num_candidates.argtypes = [vp]
    lib.pgm_num_candidates.restype = C.c_longlong
    lib.pgm_list_check_state.argtypes = [vp, vp]
    lib.pgm_late_list_checks.argtypes = [vp]
    lib.pgm_late_list_checks.restype = C.c_long
    lib.pgm_launch_pending.argtypes = [vp]
    lib.pgm_field.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.pgm_set_field_otf.argtypes = [vp, C.c_int]
    lib.pgm_field_otf.argtypes = [vp]
    lib.pgm_setup_lanes.argtypes = [vp]
    lib.pgm_set_setup_lanes.argtypes = [vp, C.c_int]
    lib.pgm_set_pme_overlap.argtypes = [vp, C.c_int]
    lib.pgm_pme_overlap.argtypes = [vp]
    lib.pgm_pme_overlap_times.argtypes = [vp, vp]
    lib.pgm_set_history.argtypes = [vp, C.c_int, vp]
    lib.pgm_freeze_history.argtypes = [vp, C.c_int]
    lib.pgm_set_timing.argtypes = [vp, C.c_int]
    lib.pgm_set_timing_phases.argtypes = [vp, C.c_uint]
    lib.pgm_num_pairs.argtypes = [vp]
    lib.pgm_num_pairs.restype = C.c_longlong
    lib.pgm_get_timing.argtypes = [vp, vp, vp, C.c_int]
    lib.pgm_backward.argtypes = [vp, vp, vp, vp, C.c_double, vp, vp, vp, vp]
    lib.pgm_backward_device.argtypes = [vp, vp, vp, vp, C.c_double, vp, vp, vp, vp, vp]
    lib.pgm_adjoint_iterations.argtypes = [vp]
    lib.pgm_adjoint_residual.argtypes = [vp]
    lib.pgm_adjoint_residual.restype = C.c_double
    lib.pgm_set_batch.argtypes = [vp, C.c_int, vp, vp]
    lib.pgm_num_systems.argtypes = [vp]
    lib.pgm_compute_batch.argtypes = [vp, vp, vp, C.c_int, vp, vp]
    lib.pgm_compute_batch_device.argtypes = [vp, vp, vp, C.c_int, vp, vp, vp]
    lib.pgm_backward_batch.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.pgm_backward_batch_device.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.pgm_autotune.argtypes = [vp, C.c_int, vp, vp, vp]
    lib.pgm_set_lanes.argtypes = [vp, C.c_int, C.c_int]
    lib.pgm_eval_erfc_f32.argtypes = [C.c_int, vp, vp]
    lib.pgm_set_gvdw.argtypes = [vp, C.c_int, C.c_double, vp, vp, vp, C.c_int, vp, vp]
    lib.pgm_set_preconditioner.argtypes = [vp, C.c_double, C.c_int]
    lib.pgm_set_candidates.argtypes = [vp, C.c_int, vp, vp, vp, vp]
    lib.pgm_wait_device.argtypes = [vp, vp]
    lib.pgm_eval_pair_kernels.argtypes = [C.c_int, C.c_int, vp, vp, C.c_double, vp]
    _ = dp
    _lib = lib
    return lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def make_settings(**kw) -> Settings:
    lib = load()
    s = Settings()
    lib.pgm_default_settings(C.byref(s))
    for k, v in kw.items():
        if k == "grid":
            for a in range(3):
                s.grid[a] = int(v[a])
        elif k == "precision":
            s.precision = {"double": PGM_DOUBLE, "mixed": PGM_MIXED}.get(v, v)
        elif k in ("periodic",):
            if not v:
                raise PgmError("the HIP engine runs periodic systems (PME) only")
        elif hasattr(s, k):
            setattr(s, k, type(getattr(s, k))(v))
        else:
            raise PgmError(f"unknown setting {k}")
    return s


class Engine:
    """pGM electrostatics on the GPU; parameters and positions in the caller's atom order."""

    def __init__(self, settings: Settings | None = None, **kw):
        self.lib = load()
        self.settings = settings if settings is not None else make_settings(**kw)
        self.h = self.lib.pgm_create(C.byref(self.settings))
        if not self.h:
            raise PgmError(self.lib.pgm_last_error(None).decode())
        self.n = 0
        self.ncov = 0
        self.nsys = 1

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.pgm_destroy(self.h)
            self.h = None

    def _check(self, rc: int):
        if rc != 0:
            raise PgmError(self.lib.pgm_last_error(self.h).decode())

    @classmethod
    def for_system(cls, system, precision="mixed", **overrides) -> "Engine":
        st = dict(system.settings)
        st.update(overrides)
        keys = ("cutoff", "beta", "grid", "order", "tol", "max_iter", "peek", "predictor", "ke", "skin", "polarizable")
        kw = {k: st[k] for k in keys if k in st}
        if "device" in st:
            kw["device"] = st["device"]
        e = cls(precision=precision, **kw)
        e.set_atoms(system.q, system.radius, system.alpha)
        e.set_covalent(system.cov_atom, system.cov_partner, system.cov_strength)
        return e

    def set_atoms(self, q, radius, alpha, global_index=None, n_global=None):
        q, radius, alpha = _f64(q), _f64(radius), _f64(alpha)
        gi = None if global_index is None else np.ascontiguousarray(global_index, np.int32)
        self.n = len(q)
        ng = n_global if n_global is not None else self.n
        self._keep = (q, radius, alpha, gi)
        self._check(self.lib.pgm_set_atoms(self.h, self.n, _ptr(q), _ptr(radius), _ptr(alpha), _ptr(gi), ng))

    def set_covalent(self, atom, partner, strength):
        a = np.ascontiguousarray(atom, np.int32)
        p = np.ascontiguousarray(partner, np.int32)
        s = _f64(strength)
        self._check(self.lib.pgm_set_covalent(self.h, len(a), _ptr(a), _ptr(p), _ptr(s)))
        self.ncov = len(a)

    def compute(self, pos, box, new_step=True, forces=True, strain=False):
        """Return (terms dict, forces (n,3) or None, strain (3,3) or None)."""
        pos = _f64(pos).reshape(-1, 3)
        if self.n and len(pos) != self.n:
            raise PgmError(f"positions for {len(pos)} atoms, engine has {self.n}")
        box = _f64(box, (3, 3))
        f = np.zeros((self.n, 3)) if forces else None
        S = np.zeros(9) if strain else None
        t = Terms()
        self._check(self.lib.pgm_compute(self.h, _ptr(pos), _ptr(box), int(bool(new_step)), _ptr(f), _ptr(S), C.byref(t)))
        return t.as_dict(), f, (S.reshape(3, 3) if strain else None)

    def compute_f32(self, pos, box, new_step=True):
        pos = np.ascontiguousarray(pos, np.float32).reshape(self.n, 3)
        box = _f64(box, (3, 3))
        t = Terms()
        self._check(self.lib.pgm_compute_f32(self.h, _ptr(pos), _ptr(box), int(bool(new_step)), None, C.byref(t)))
        f = np.zeros((self.n, 3))
        self._check(self.lib.pgm_forces(self.h, _ptr(f)))
        return t.as_dict(), f

    def compute_device_f64(self, pos_ptr: int, box, new_step=True, f_ptr: int | None = None, stream: int = 0):
        box = _f64(box, (3, 3))
        t = Terms()
        self._check(self.lib.pgm_compute_device_f64(self.h, pos_ptr, _ptr(box), int(bool(new_step)), f_ptr, None,
                                                    stream, C.byref(t)))  # fmt: skip
        return t.as_dict()

    def compute_device_f64_noalloc(self, pos_ptr: int, box, f_ptr: int | None = None, stream: int = 0):
        """pgm_compute_device_f64_noalloc: one evaluation with no device allocation (static scratch)."""
        box = _f64(box, (3, 3))
        t = Terms()
        self._check(self.lib.pgm_compute_device_f64_noalloc(self.h, pos_ptr, _ptr(box), f_ptr, None, stream,
                                                            C.byref(t)))  # fmt: skip
        return t.as_dict()

    def field(self, pos, box, q, d):
        """pgm_field: the field (n, 3) of charges q and dipoles d at every atom at these positions."""
        pos = _f64(pos).reshape(-1, 3)
        q, d = _f64(q), _f64(d).reshape(-1, 3)
        out = np.zeros((self.n, 3))
        self._check(self.lib.pgm_field(self.h, _ptr(pos), _ptr(_f64(box, (3, 3))), _ptr(q), _ptr(d), _ptr(out)))
        return out

    def compute_device(self, posq_ptr: int, box, new_step, fixed_ptr: int, stride: int, scale: float, stream: int = 0):
        box = _f64(box, (3, 3))
        t = Terms()
        rc = self.lib.pgm_compute_device(
            self.h, posq_ptr, _ptr(box), int(bool(new_step)), fixed_ptr, stride, scale, stream, C.byref(t)
        )
        self._check(rc)
        return t.as_dict()

    def launch_device(self, posq_ptr: int, box, new_step, fixed_ptr: int, stride: int, scale: float, stream: int = 0):
        """pgm_launch_device: compute_device in two parts; finish_device or wait_device follows."""
        box = _f64(box, (3, 3))
        self._check(self.lib.pgm_launch_device(self.h, posq_ptr, _ptr(box), int(bool(new_step)), fixed_ptr, stride,
                                               scale, stream))  # fmt: skip

    def finish_device(self):
        """pgm_finish_device: (terms, valid) of the last launch."""
        t = Terms()
        valid = C.c_int(0)
        self._check(self.lib.pgm_finish_device(self.h, C.byref(t), C.byref(valid)))
        return t.as_dict(), bool(valid.value)

    def wait_device(self, stream: int):
        """pgm_wait_device: the stream waits on the GPU for the last launch's forces."""
        self._check(self.lib.pgm_wait_device(self.h, stream))

    def launch_pending(self) -> bool:
        """pgm_launch_pending: the GPU work of the launch is still running."""
        return bool(self.lib.pgm_launch_pending(self.h))

    def late_list_checks(self) -> int:
        return self.lib.pgm_late_list_checks(self.h)

    def list_check_state(self) -> dict:
        """pgm_list_check_state: the inputs of the next launch's list prediction (test hook)."""
        out = np.zeros(5)
        self._check(self.lib.pgm_list_check_state(self.h, _ptr(out)))
        return {"displacement": float(out[0]), "last_move": float(out[1]), "skin": float(out[2]),
                "repeats": int(out[3]), "sync_short_solves": int(out[4])}

    def backward(self, pos, box, dl_dforces=None, dl_denergy=0.0) -> dict:
        """Gradients of a loss L(E, F) with respect to the per-atom parameters (pgm_backward).

        dl_dforces: dL/dF (n,3) or None; dl_denergy: dL/dE.  Returns {"q", "radius", "alpha"} (n,) and
        {"cov_strength"} (ncov,) in the orders of set_atoms / set_covalent.
        """
        pos = _f64(pos).reshape(-1, 3)
        if len(pos) != self.n:
            raise PgmError(f"positions for {len(pos)} atoms, engine has {self.n}")
        box = _f64(box, (3, 3))
        g = None if dl_dforces is None else _f64(dl_dforces).reshape(self.n, 3)
        gq, gr, ga = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        gc = np.zeros(max(self.ncov, 1))
        rc = self.lib.pgm_backward(self.h, _ptr(pos), _ptr(box), _ptr(g), float(dl_denergy), _ptr(gq), _ptr(gr),
                                   _ptr(ga), _ptr(gc))  # fmt: skip
        self._check(rc)
        return {"q": gq, "radius": gr, "alpha": ga, "cov_strength": gc[: self.ncov]}

    def backward_device(self, pos_ptr: int, box, g_ptr: int | None, dl_denergy: float, gq_ptr, gr_ptr, ga_ptr, gc_ptr,
                        stream: int = 0):  # fmt: skip
        box = _f64(box, (3, 3))
        rc = self.lib.pgm_backward_device(self.h, pos_ptr, _ptr(box), g_ptr, float(dl_denergy), gq_ptr, gr_ptr, ga_ptr,
                                          gc_ptr, stream)  # fmt: skip
        self._check(rc)

    def set_gvdw(self, repulsion: int, rvdw: float, sqrt_a, sqrt_c6, b, skip_i=(), skip_j=()):
        """GVDW by the combination rule within rvdw, except the skipped pairs (pgm_set_gvdw); rvdw 0: off."""
        sa, sc, bb = (_f64(v).reshape(-1) for v in (sqrt_a, sqrt_c6, b))
        si = np.ascontiguousarray(skip_i, np.int32)
        sj = np.ascontiguousarray(skip_j, np.int32)
        if len(si) != len(sj):
            raise PgmError("skip_i and skip_j differ in length")
        self._check(self.lib.pgm_set_gvdw(self.h, int(repulsion), float(rvdw), _ptr(sa), _ptr(sc), _ptr(bb), len(si),
                                          _ptr(si) if len(si) else None, _ptr(sj) if len(sj) else None))  # fmt: skip

    def set_candidates(self, seg_i, seg_shift, seg_begin, cand_j):
        """The caller's pair list (pgm_set_candidates): segment s = pairs (seg_i[s], cand_j[k]),
        k in [seg_begin[s], seg_begin[s+1]), image x = r_i + seg_shift[s] . box - r_j."""
        si = np.ascontiguousarray(seg_i, np.int32)
        sh = np.ascontiguousarray(seg_shift, np.int32).reshape(-1)
        sb = np.ascontiguousarray(seg_begin, np.int32)
        cj = np.ascontiguousarray(cand_j, np.int32)
        if len(sb) != len(si) + 1 or len(sh) != 3 * len(si):
            raise ValueError("seg_begin needs nseg + 1 entries and seg_shift 3 per segment")
        self._check(self.lib.pgm_set_candidates(self.h, len(si), _ptr(si) if len(si) else None,
                                                _ptr(sh) if len(sh) else None, _ptr(sb), _ptr(cj) if len(cj) else None))

    def set_preconditioner(self, radius: float, jacobi_self: bool = False):
        """CG preconditioner (pgm_set_preconditioner): short-range M = D + D T_s D within radius [nm] (0: Jacobi);
        jacobi_self: Jacobi with the self term, alpha/(1 - alpha c_self)."""
        self._check(self.lib.pgm_set_preconditioner(self.h, float(radius), int(bool(jacobi_self))))

    def set_batch(self, atom_offsets, grids=None):
        """Partition the atoms into independent systems (pgm_set_batch); grids (nsys, 3) or None."""
        off = np.ascontiguousarray(atom_offsets, np.int32)
        g = None if grids is None else np.ascontiguousarray(np.asarray(grids).reshape(-1, 3), np.int32)
        self._check(self.lib.pgm_set_batch(self.h, len(off) - 1, _ptr(off), _ptr(g)))
        self.nsys = len(off) - 1

    def num_systems(self) -> int:
        return self.lib.pgm_num_systems(self.h)

    def compute_batch(self, pos, boxes, new_step=True, forces=True):
        """Return ([terms dict per system], forces (n,3) or None)."""
        pos = _f64(pos).reshape(-1, 3)
        if len(pos) != self.n:
            raise PgmError(f"positions for {len(pos)} atoms, engine has {self.n}")
        nsys = self.num_systems()
        boxes = _f64(boxes).reshape(nsys, 3, 3)
        f = np.zeros((self.n, 3)) if forces else None
        terms = (Terms * nsys)()
        rc = self.lib.pgm_compute_batch(self.h, _ptr(pos), _ptr(boxes), int(bool(new_step)), _ptr(f), terms)
        self._check(rc)
        return [t.as_dict() for t in terms], f

    def compute_batch_device(self, pos_ptr: int, boxes, new_step=True, f_ptr: int | None = None, stream: int = 0):
        nsys = self.num_systems()
        boxes = _f64(boxes).reshape(nsys, 3, 3)
        terms = (Terms * nsys)()
        rc = self.lib.pgm_compute_batch_device(self.h, pos_ptr, _ptr(boxes), int(bool(new_step)), f_ptr, stream, terms)
        self._check(rc)
        return [t.as_dict() for t in terms]

    def backward_batch(self, pos, boxes, dl_dforces=None, dl_denergy=None) -> dict:
        """pgm_backward_batch: dl_denergy per system (or None)."""
        pos = _f64(pos).reshape(-1, 3)
        nsys = self.num_systems()
        boxes = _f64(boxes).reshape(nsys, 3, 3)
        g = None if dl_dforces is None else _f64(dl_dforces).reshape(self.n, 3)
        we = None if dl_denergy is None else _f64(np.broadcast_to(np.asarray(dl_denergy, float), (nsys,)))
        gq, gr, ga = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        gc = np.zeros(max(self.ncov, 1))
        rc = self.lib.pgm_backward_batch(self.h, _ptr(pos), _ptr(boxes), _ptr(g), _ptr(we), _ptr(gq), _ptr(gr),
                                         _ptr(ga), _ptr(gc))  # fmt: skip
        self._check(rc)
        return {"q": gq, "radius": gr, "alpha": ga, "cov_strength": gc[: self.ncov]}

    def backward_batch_device(self, pos_ptr: int, boxes, g_ptr, dl_denergy, gq_ptr, gr_ptr, ga_ptr, gc_ptr,
                              stream: int = 0):  # fmt: skip
        nsys = self.num_systems()
        boxes = _f64(boxes).reshape(nsys, 3, 3)
        we = None if dl_denergy is None else _f64(np.broadcast_to(np.asarray(dl_denergy, float), (nsys,)))
        rc = self.lib.pgm_backward_batch_device(self.h, pos_ptr, _ptr(boxes), g_ptr, _ptr(we), gq_ptr, gr_ptr, ga_ptr,
                                                gc_ptr, stream)  # fmt: skip
        self._check(rc)

    def autotune(self, reps: int = 3):
        """((lanes, fused) chosen, {(lanes, fused): ms per evaluation in the pair passes}) (pgm_autotune)."""
        lanes, fused = C.c_int(0), C.c_int(0)
        ms = np.zeros(6)
        self._check(self.lib.pgm_autotune(self.h, int(reps), C.byref(lanes), C.byref(fused), _ptr(ms)))
        table = {(L, f): float(ms[3 * (0 if f else 1) + k]) for f in (True, False) for k, L in enumerate((16, 32, 64))}
        return (lanes.value, bool(fused.value)), table

    def set_lanes(self, lanes: int, fused: bool = True):
        self._check(self.lib.pgm_set_lanes(self.h, int(lanes), int(bool(fused))))

    def setup_lanes(self) -> int:
        """Lanes per atom of the pair setup pass (pgm_setup_lanes)."""
        return self.lib.pgm_setup_lanes(self.h)

    def set_setup_lanes(self, lanes: int):
        self._check(self.lib.pgm_set_setup_lanes(self.h, int(lanes)))

    def set_field_otf(self, on: bool = True):
        """FP32 field pass with its kernels evaluated on the fly (pgm_set_field_otf)."""
        self._check(self.lib.pgm_set_field_otf(self.h, int(bool(on))))

    def field_otf(self) -> bool:
        return self.lib.pgm_field_otf(self.h) == 1

    def set_pme_overlap(self, on: bool = True):
        """PME solves beside the pair passes on a stream of their own (pgm_set_pme_overlap)."""
        self._check(self.lib.pgm_set_pme_overlap(self.h, int(bool(on))))

    def pme_overlap(self) -> bool:
        return self.lib.pgm_pme_overlap(self.h) == 1

    def pme_overlap_times(self) -> tuple:
        """The autotuner's ms per CG product: (PME after the field pass, PME beside it)."""
        ms = np.zeros(2)
        self._check(self.lib.pgm_pme_overlap_times(self.h, _ptr(ms)))
        return float(ms[0]), float(ms[1])

    def adjoint_iterations(self) -> int:
        return self.lib.pgm_adjoint_iterations(self.h)

    def adjoint_residual(self) -> float:
        return self.lib.pgm_adjoint_residual(self.h)

    def iterations(self) -> int:
        return self.lib.pgm_iterations(self.h)

    def residual(self) -> float:
        return self.lib.pgm_residual(self.h)

    def induced_dipoles(self):
        m = np.zeros((self.n, 3))
        self._check(self.lib.pgm_induced_dipoles(self.h, _ptr(m)))
        return m

    def permanent_dipoles(self):
        m = np.zeros((self.n, 3))
        self._check(self.lib.pgm_permanent_dipoles(self.h, _ptr(m)))
        return m

    def reset_history(self):
        self._check(self.lib.pgm_reset_history(self.h))

    def set_history(self, sets):
        sets = np.ascontiguousarray(np.asarray(sets, np.float64))
        self._check(self.lib.pgm_set_history(self.h, len(sets), _ptr(sets)))

    def freeze_history(self, frozen=True):
        self._check(self.lib.pgm_freeze_history(self.h, int(bool(frozen))))

    def set_timing(self, on=True, phases=None):
        """Time every phase (phases None) or only the named ones (pgm_set_timing_phases)."""
        if phases is None:
            self._check(self.lib.pgm_set_timing(self.h, int(bool(on))))
        else:
            mask = sum(1 << PHASES.index(p) for p in phases) if on else 0
            self._check(self.lib.pgm_set_timing_phases(self.h, mask))

    def timing(self) -> dict:
        """{phase: (total ms, launches counted)} since set_timing."""
        ms = np.zeros(len(PHASES))
        cnt = np.zeros(len(PHASES), dtype=np.int64)
        self._check(self.lib.pgm_get_timing(self.h, _ptr(ms), _ptr(cnt), len(PHASES)))
        return {p: (float(ms[k]), int(cnt[k])) for k, p in enumerate(PHASES)}

    def num_pairs(self) -> int:
        return self.lib.pgm_num_pairs(self.h)

    def list_builds(self) -> int:
        return self.lib.pgm_list_builds(self.h)

    def num_candidates(self) -> int:
        return self.lib.pgm_num_candidates(self.h)


def eval_erfc_f32(x) -> np.ndarray:
    """erfc as the FP32 pair kernels evaluate it on the device (pgm_eval_erfc_f32)."""
    lib = load() if _lib is None else _lib
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    if lib.pgm_eval_erfc_f32(len(x), _ptr(x), _ptr(y)) != 0:
        raise PgmError(lib.pgm_last_error(None).decode())
    return y


def eval_pair_kernels(precision: str, rg_sum, r2, beta: float) -> np.ndarray:
    """(n, 6): G0..G3 of the device pair kernels, then G0, G3 as the FP32 passes rebuild them (pgm_eval_pair_kernels)."""
    lib = load() if _lib is None else _lib
    rg_sum = np.ascontiguousarray(rg_sum, dtype=np.float64)
    r2 = np.ascontiguousarray(r2, dtype=np.float64)
    out = np.empty((len(r2), 6))
    prec = {"double": 0, "mixed": 1}[precision]
    if lib.pgm_eval_pair_kernels(prec, len(r2), _ptr(rg_sum), _ptr(r2), float(beta), _ptr(out)) != 0:
        raise PgmError(lib.pgm_last_error(None).decode())
    return out


def total_energy(terms: dict) -> float:
    return sum(terms.values())


def rel_rms(f, fref) -> float:
    """sqrt(mean(dF^2)) / sqrt(mean(F_ref^2)) over all components (tests/xplat/omm_vs_amber.py:39-41)."""
    f, fref = np.asarray(f), np.asarray(fref)
    return float(np.sqrt(((f - fref) ** 2).mean()) / np.sqrt((fref**2).mean()))
