"""The coefficients of erfcxFast in device.hpp: exp(x^2) erfc(x) = t Q(t), t = 1/(1 + x/2), x in [0, 4.5], with
Q of degree 8 fitted for the least maximum relative error by iteratively reweighted least squares (Lawson's
method) on Chebyshev nodes in t.  Prints the FP32 coefficients, highest degree first (Horner order), and the
fit's maximum relative error.

    python pgm-md_amd/csrc/erfcx_fit.py
"""

import numpy as np
from scipy.special import erfcx

DEG, XMAX, NODES, ITERS = 8, 4.5, 4000, 60


def fit():
    t0 = 1.0 / (1.0 + XMAX / 2.0)
    k = np.arange(NODES)
    t = 0.5 * (1.0 + t0) + 0.5 * (1.0 - t0) * np.cos(np.pi * (k + 0.5) / NODES)
    f = erfcx(2.0 * (1.0 / t - 1.0)) / t
    w = np.ones_like(f)
    best = None
    V = np.vander(t, DEG + 1)
    for _ in range(ITERS):
        c, *_ = np.linalg.lstsq(V * (w / f)[:, None], w, rcond=None)
        rel = (V @ c - f) / f
        m = np.abs(rel).max()
        if best is None or m < best[0]:
            best = (m, c.copy())
        w = w * (1.0 + 5.0 * np.abs(rel) / m)
    return best


if __name__ == "__main__":
    err, c = fit()
    print(", ".join(f"{v:.9e}" for v in c.astype(np.float32)))
    print(f"max relative error of the fit: {err:.2e}")
