"""The C-ABI library: it loads, and it exports every entry point include/pgm_hip.h declares (no GPU needed).
Settings validation runs before any device call, so bad settings are rejected here too."""

import ctypes as C
import os
import re

import pytest

from pgmhip import engine as E

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "pgm_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pgm_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = E.load()
    names = declared()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    assert set(E.EXPORTS) <= set(names)


def test_default_settings_mirror_the_reference_defaults():
    s = E.make_settings()
    # pgmcuda::Settings defaults (core/cuda/pgm_cuda.h:25-39) with pgmcore's Coulomb constant
    assert (s.cutoff, s.beta, list(s.grid), s.order) == (0.9, 3.5, [36, 36, 36], 6)
    assert (s.tol, s.max_iter, s.peek, s.predictor, s.skin) == (1e-5, 100, 0.65, 2, 0.15)
    assert s.ke == 138.935456 and s.precision == E.PGM_MIXED and s.polarizable == 1


@pytest.mark.parametrize(
    "bad",
    [dict(order=3), dict(order=9), dict(grid=(32, 32, 4)), dict(grid=(32, 32, 143)), dict(tol=0.0), dict(beta=-1.0),
     dict(cutoff=0.0), dict(predictor=3), dict(skin=-0.1), dict(precision=7), dict(max_iter=-1)],
)  # fmt: skip
def test_invalid_settings_are_rejected(bad):
    lib = E.load()
    s = E.make_settings(**bad)
    assert lib.pgm_create(C.byref(s)) is None
    assert b"pgm_hip:" in lib.pgm_last_error(None)


def test_null_handle_calls_fail_cleanly():
    lib = E.load()
    assert lib.pgm_set_atoms(None, 1, None, None, None, None, 1) != 0
    assert lib.pgm_iterations(None) == -1
    assert lib.pgm_num_pairs(None) == -1


def test_device_math_hooks_check_arguments_first():
    """The device-math test hooks reject bad arguments (message in pgm_last_error(NULL)) and return at once for
    empty input, before any device call."""
    lib = E.load()
    assert lib.pgm_eval_pair_kernels(7, 1, None, None, 1.0, None) != 0
    assert "precision" in lib.pgm_last_error(None).decode()
    assert lib.pgm_eval_erfc_f32(-1, None, None) != 0
    assert "n >= 0" in lib.pgm_last_error(None).decode()
    assert lib.pgm_eval_erfc_f32(0, None, None) == 0


def test_only_the_c_abi_is_exported():
    """No C++ symbol leaves the library: a caller's own classes (e.g. the C++ facade's pgmhip::Engine) cannot
    interpose on the engine's internals."""
    import subprocess

    out = subprocess.run(["nm", "-D", "--defined-only", E.LIB_PATH], capture_output=True, text=True, check=True).stdout
    names = [ln.split()[-1] for ln in out.splitlines() if ln.strip()]
    exported = [n for n in names if not n.startswith("__hip") and n not in ("_init", "_fini")]
    assert exported and all(n.startswith("pgm_") for n in exported), [n for n in exported if not n.startswith("pgm_")]
