"""Test infrastructure only: ctypes wrapper of the FP64 CPU oracle (liborc.so) and of the recipe-built
reference library (_ref/libpgm_ref.so), which export the same C ABI (pgm_oracle.h).

Only tests/, the smoke check and the bench's CPU-baseline leg import this module; the engine never does.
The wrapper mirrors pgmcore::Engine (core/pgm_core.h:47-115).
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PORT_LIB = os.path.join(HERE, "liborc.so")
REF_LIB = os.path.join(HERE, "_ref", "libpgm_ref.so")


class Settings(C.Structure):
    """orc_settings = pgmcore::Settings (core/pgm_core.h:22-34)."""

    _fields_ = [
        ("periodic", C.c_int),
        ("cutoff", C.c_double),
        ("beta", C.c_double),
        ("grid", C.c_int * 3),
        ("order", C.c_int),
        ("polarizable", C.c_int),
        ("tol", C.c_double),
        ("maxIter", C.c_int),
        ("peek", C.c_double),
        ("predictor", C.c_int),
        ("ke", C.c_double),
    ]


_DP = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_IP = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_libs: dict[str, C.CDLL] = {}


def load(kind: str = "port") -> C.CDLL:
    """Load the oracle library: "port" (this restatement) or "reference" (oracle/_ref)."""
    if kind in _libs:
        return _libs[kind]
    path = PORT_LIB if kind == "port" else REF_LIB
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built (make -C oracle)")
    lib = C.CDLL(path)
    lib.orc_last_error.restype = C.c_char_p
    lib.orc_kind.restype = C.c_char_p
    lib.orc_create.restype = C.c_void_p
    lib.orc_create.argtypes = [C.POINTER(Settings), C.c_int, _DP, _DP, _DP, C.c_int, _IP, _IP, _DP]
    lib.orc_destroy.argtypes = [C.c_void_p]
    lib.orc_compute.argtypes = [C.c_void_p, _DP, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]
    lib.orc_terms.argtypes = [C.c_void_p, _DP]
    lib.orc_induced.argtypes = [C.c_void_p, _DP]
    lib.orc_permanent.argtypes = [C.c_void_p, _DP]
    lib.orc_iterations.argtypes = [C.c_void_p]
    lib.orc_set_preconditioner.argtypes = [C.c_void_p, C.c_double, C.c_int]
    lib.orc_precondition.argtypes = [C.c_void_p, _DP, _DP]
    lib.orc_preconditioner_on.argtypes = [C.c_void_p]
    lib.orc_residual.argtypes = [C.c_void_p]
    lib.orc_residual.restype = C.c_double
    lib.orc_strain.argtypes = [C.c_void_p, _DP]
    lib.orc_set_history_key.argtypes = [C.c_void_p, C.c_longlong]
    lib.orc_reset_history.argtypes = [C.c_void_p]
    lib.orc_push_history.argtypes = [C.c_void_p, _DP]
    lib.orc_set_topology.argtypes = [C.c_void_p, _DP, _DP, _DP, C.c_int, _IP, _IP, _DP]
    lib.orc_set_geometry.argtypes = [C.c_void_p, _DP, C.c_void_p]
    lib.orc_field.argtypes = [C.c_void_p, _DP, _DP, _DP]
    lib.orc_num_pairs.argtypes = [C.c_void_p]
    lib.orc_num_pairs.restype = C.c_longlong
    for name in ("orc_erf_kernels", "orc_erfc_kernels"):
        getattr(lib, name).argtypes = [C.c_double, C.c_double, C.c_int, _DP]
    lib.orc_pair_kernels.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, _DP]
    lib.orc_bspline_moduli.argtypes = [C.c_int, C.c_int, _DP]
    lib.orc_gvdw_g.argtypes = [C.c_double, _DP, _DP]
    lib.orc_gvdw_pair.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _DP, _DP]
    lib.orc_gvdw_sum.argtypes = [C.c_int, _DP, _DP, _DP, _DP, _DP, _DP, C.c_int, C.c_double, C.c_int, _IP, _IP,
                                 _DP, C.c_void_p, C.c_void_p]
    lib.orc_ewald_coefficient.argtypes = [C.c_double, C.c_double]
    lib.orc_ewald_coefficient.restype = C.c_double
    lib.orc_fft3d_r2c.argtypes = [C.c_int, C.c_int, C.c_int, _DP, _DP]
    lib.orc_fft3d_c2r.argtypes = [C.c_int, C.c_int, C.c_int, _DP, _DP]
    _libs[kind] = lib
    return lib


def _check(lib, rc):
    if rc != 0:
        raise RuntimeError(lib.orc_last_error().decode())


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


class Oracle:
    """pgmcore::Engine behind the C ABI: compute(), field(), dipoles, iterations, strain, history."""

    def __init__(self, system, kind: str = "port", **overrides):
        self.lib = load(kind)
        self.kind = kind
        st = dict(system.settings)
        st.update(overrides)
        s = Settings()
        s.periodic = int(st.get("periodic", 1))
        s.cutoff = st["cutoff"]
        s.beta = st["beta"]
        for a in range(3):
            s.grid[a] = int(st["grid"][a])
        s.order = int(st["order"])
        s.polarizable = int(st.get("polarizable", 1))
        s.tol = st["tol"]
        s.maxIter = int(st.get("max_iter", 100))
        s.peek = st.get("peek", 0.65)
        s.predictor = int(st.get("predictor", 2))
        s.ke = st.get("ke", 138.935456)
        self.settings = s
        self.n = len(system.q)
        self.h = self.lib.orc_create(
            C.byref(s), self.n, _f64(system.q), _f64(system.radius), _f64(system.alpha),
            len(system.cov_atom), np.ascontiguousarray(system.cov_atom, np.int32),
            np.ascontiguousarray(system.cov_partner, np.int32), _f64(system.cov_strength),
        )  # fmt: skip
        if not self.h:
            raise RuntimeError(self.lib.orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def compute(self, pos, box, forces: bool = True):
        """Return (energy, forces or None); forces start from zero."""
        pos = _f64(pos, (self.n, 3))
        box = _f64(box, (3, 3)) if box is not None else None
        f = np.zeros((self.n, 3)) if forces else None
        e = C.c_double()
        rc = self.lib.orc_compute(
            self.h, pos, box.ctypes.data if box is not None else None, f.ctypes.data if forces else None, C.byref(e)
        )
        _check(self.lib, rc)
        return e.value, f

    def terms(self):
        t = np.zeros(5)
        _check(self.lib, self.lib.orc_terms(self.h, t))
        return dict(zip(("direct", "reciprocal", "self", "background", "polarization"), t))

    def induced(self):
        m = np.zeros((self.n, 3))
        _check(self.lib, self.lib.orc_induced(self.h, m))
        return m

    def permanent(self):
        m = np.zeros((self.n, 3))
        _check(self.lib, self.lib.orc_permanent(self.h, m))
        return m

    def iterations(self) -> int:
        return self.lib.orc_iterations(self.h)

    def residual(self) -> float:
        return self.lib.orc_residual(self.h)

    def strain(self):
        s = np.zeros(9)
        _check(self.lib, self.lib.orc_strain(self.h, s))
        return s.reshape(3, 3)

    def set_history_key(self, key: int):
        _check(self.lib, self.lib.orc_set_history_key(self.h, int(key)))

    def reset_history(self):
        _check(self.lib, self.lib.orc_reset_history(self.h))

    def push_history(self, mu):
        """Push one prior solution (oldest first when pushing several)."""
        _check(self.lib, self.lib.orc_push_history(self.h, _f64(mu, (self.n, 3))))

    def set_topology(self, q, radius, alpha, cov_atom, cov_partner, cov_strength):
        rc = self.lib.orc_set_topology(
            self.h, _f64(q), _f64(radius), _f64(alpha), len(cov_atom), np.ascontiguousarray(cov_atom, np.int32),
            np.ascontiguousarray(cov_partner, np.int32), _f64(cov_strength),
        )  # fmt: skip
        _check(self.lib, rc)

    def set_geometry(self, pos, box):
        box = _f64(box, (3, 3)) if box is not None else None
        rc = self.lib.orc_set_geometry(self.h, _f64(pos, (self.n, 3)), box.ctypes.data if box is not None else None)
        _check(self.lib, rc)

    def field(self, q, d):
        E = np.zeros((self.n, 3))
        _check(self.lib, self.lib.orc_field(self.h, _f64(q), _f64(d, (self.n, 3)), E))
        return E

    def num_pairs(self) -> int:
        return self.lib.orc_num_pairs(self.h)

    def set_preconditioner(self, radius: float, jacobi_self: bool = False):
        """CG preconditioner: short-range M = D + D T_s D within radius (0: Jacobi); Jacobi with the self term."""
        _check(self.lib, self.lib.orc_set_preconditioner(self.h, float(radius), int(bool(jacobi_self))))

    def precondition(self, r):
        """z = M r for a 3n vector r (after set_geometry or compute)."""
        z = np.zeros((self.n, 3))
        _check(self.lib, self.lib.orc_precondition(self.h, _f64(r, (self.n, 3)), z))
        return z

    def preconditioner_on(self) -> bool:
        return self.lib.orc_preconditioner_on(self.h) == 1


def pair_kernels(a, beta, r, nmax=4, kind="port"):
    lib = load(kind)
    G = np.zeros(4)
    _check(lib, lib.orc_pair_kernels(a, beta, r, nmax, G))
    return G[:nmax]


def erf_kernels(a, r, nmax=4, kind="port"):
    lib = load(kind)
    B = np.zeros(4)
    _check(lib, lib.orc_erf_kernels(a, r, nmax, B))
    return B[:nmax]


def erfc_kernels(a, r, nmax=4, kind="port"):
    lib = load(kind)
    B = np.zeros(4)
    _check(lib, lib.orc_erfc_kernels(a, r, nmax, B))
    return B[:nmax]


def gvdw_g(y, kind="port"):
    """(G, H) of GVDW at y (core/gvdw_kernel.h:30-55)."""
    lib = load(kind)
    G, H = np.zeros(1), np.zeros(1)
    _check(lib, lib.orc_gvdw_g(float(y), G, H))
    return float(G[0]), float(H[0])


def gvdw_pair(r, bij, A, C6, B, rep, kind="port"):
    """(energy, dEdR = -(1/r) dU/dr) of one GVDW pair (core/gvdw_kernel.h:58-77)."""
    lib = load(kind)
    e, d = np.zeros(1), np.zeros(1)
    _check(lib, lib.orc_gvdw_pair(float(r), float(bij), float(A), float(C6), float(B), int(rep), e, d))
    return float(e[0]), float(d[0])


def gvdw_sum(pos, box, radius, sqrt_a, sqrt_c6, b, rep, rvdw, skip_i=(), skip_j=(), kind="port"):
    """GVDW by the combination rule over all minimum-image pairs within rvdw except the skipped ones, brute
    force: (energy [kJ/mol], forces (n, 3), strain (3, 3) = dE/d eps)."""
    lib = load(kind)
    pos = _f64(pos).reshape(-1, 3)
    n = len(pos)
    si = np.ascontiguousarray(skip_i, np.int32)
    sj = np.ascontiguousarray(skip_j, np.int32)
    e, f, S = np.zeros(1), np.zeros((n, 3)), np.zeros(9)
    _check(lib, lib.orc_gvdw_sum(n, pos, _f64(box, (3, 3)), _f64(radius), _f64(sqrt_a), _f64(sqrt_c6), _f64(b),
                                 int(rep), float(rvdw), len(si), si, sj, e, f.ctypes.data, S.ctypes.data))  # fmt: skip
    return float(e[0]), f, S.reshape(3, 3)


def bspline_moduli(K, p, kind="port"):
    lib = load(kind)
    out = np.zeros(K)
    _check(lib, lib.orc_bspline_moduli(K, p, out))
    return out


def ewald_coefficient(rc, rtol, kind="port"):
    return load(kind).orc_ewald_coefficient(rc, rtol)


def fft3d_r2c(x, kind="port"):
    lib = load(kind)
    K1, K2, K3 = x.shape
    out = np.zeros((K1, K2, K3 // 2 + 1, 2))
    _check(lib, lib.orc_fft3d_r2c(K1, K2, K3, _f64(x), out))
    return out[..., 0] + 1j * out[..., 1]


def fft3d_c2r(F, K3, kind="port"):
    lib = load(kind)
    K1, K2, _ = F.shape
    inp = np.ascontiguousarray(np.stack([F.real, F.imag], axis=-1), np.float64)
    out = np.zeros((K1, K2, K3))
    _check(lib, lib.orc_fft3d_c2r(K1, K2, K3, inp, out))
    return out
