"""The allocation-free entry point (pgm_compute_device_f64_noalloc): every evaluation buffer in the library's
static device scratch.  Its results equal pgm_compute_device_f64's on a fresh engine (same alpha E guess), on two
calls in a row, the second with a smaller system, and the engine's ordinary entry points still work after it."""

import numpy as np
import pytest

from pgmhip import engine as E
from pgmhip import systems as S

pytestmark = pytest.mark.gpu


def _small_box():
    """256 waters at the density of the 512-water box, 24^3 PME."""
    L = 2.48 * 0.5 ** (1.0 / 3.0)
    return S.water_box(256, L, 7, 7, 0.03, "water256", S.base_settings(grid=(24, 24, 24), tol=1e-6))


def _run(torch, eng, sy, noalloc):
    dev = torch.device("cuda:0")
    pos = torch.from_numpy(sy.pos.copy()).to(dev)
    f = torch.zeros((sy.n, 3), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    if noalloc:
        t = eng.compute_device_f64_noalloc(pos.data_ptr(), sy.box, f.data_ptr(), stream)
    else:
        t = eng.compute_device_f64(pos.data_ptr(), sy.box, True, f.data_ptr(), stream)
    torch.cuda.synchronize()
    return t, f.cpu().numpy(), eng.iterations()


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_noalloc_equals_compute_device_f64(precision):
    torch = pytest.importorskip("torch")
    for sy in (S.config0(), _small_box()):   # the second call with a smaller input
        eng = E.Engine.for_system(sy, precision=precision)
        ref = E.Engine.for_system(sy, precision=precision)
        t, f, it = _run(torch, eng, sy, True)
        tr, fr, itr = _run(torch, ref, sy, False)
        assert it == itr
        assert t == tr
        np.testing.assert_array_equal(f, fr)
        # the same engine again (nothing kept: the same alpha E solve), then its allocating path
        t2, f2, _ = _run(torch, eng, sy, True)
        assert t2 == tr
        np.testing.assert_array_equal(f2, fr)
        t3, f3, _ = _run(torch, eng, sy, False)
        assert t3 == tr
        np.testing.assert_array_equal(f3, fr)


def test_noalloc_rejects_large_systems():
    torch = pytest.importorskip("torch")
    sy = S.config1()   # 8,550 atoms: above the static scratch's 4,096
    eng = E.Engine.for_system(sy, precision="mixed")
    with pytest.raises(E.PgmError, match="at most 4096 atoms"):
        _run(torch, eng, sy, True)
    t, f, _ = _run(torch, eng, sy, False)   # the engine still works
    assert np.isfinite(f).all()
