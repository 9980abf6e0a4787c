"""Derive the polynomial coefficients used by the walk's own FP32 log and sin/cos.

Method: discrete minimax (Chebyshev) approximation solved as a linear program
over a dense grid, then coefficients rounded to binary32.  The functions are

  log1p(g) ~= g + g^2 * Q(g)            g in [sqrt(1/2) - 1, sqrt(2) - 1]
  sin(2*pi*r) ~= r * S(r^2)             r in [-1/8, 1/8]
  cos(2*pi*r) ~= C(r^2), C(0) = 1       r in [-1/8, 1/8]

each minimising the maximum relative error of the final function value.  The
printed hex-float literals are pasted into tmc_math.h (device/host product
code); the exhaustive error of the binary32 evaluation is checked by the tests.

Run: python tiny_mc_amd/csrc/gen_poly.py
"""
import numpy as np
from scipy.optimize import linprog


def minimax(basis, target, weight, npts):
    """min E s.t. |weight*(basis @ c - target)| <= E over the grid."""
    # HiGHS feasibility tolerances are absolute (~1e-7): scale the residuals
    # so that the optimum error is O(1) and the solution is not tolerance-limited.
    scale = 1e9
    A = basis * (weight * scale)[:, None]
    b = target * weight * scale
    m, n = A.shape
    # variables: c (n), E (1); minimise E
    cost = np.zeros(n + 1)
    cost[-1] = 1.0
    A_ub = np.vstack([np.hstack([A, -np.ones((m, 1))]), np.hstack([-A, -np.ones((m, 1))])])
    b_ub = np.concatenate([b, -b])
    res = linprog(cost, A_ub=A_ub, b_ub=b_ub, bounds=[(None, None)] * (n + 1), method="highs")
    assert res.success, res.message
    return res.x[:n], res.x[n] / scale


def hexf(x):
    return float(np.float32(x)).hex()


def fit_log(deg):
    lo, hi = np.sqrt(0.5) - 1.0, np.sqrt(2.0) - 1.0
    g = np.cos(np.linspace(np.pi, 0, 4001)) * (hi - lo) / 2 + (hi + lo) / 2
    g = g[np.abs(g) > 1e-6]
    qt = (np.log1p(g) - g) / g**2
    basis = np.vstack([g**i for i in range(deg + 1)]).T
    weight = np.abs(g**2 / np.log1p(g))
    c, e = minimax(basis, qt, weight, len(g))
    return c, e


def fit_sin(deg):
    r = np.linspace(1e-6, 0.125, 4001)
    x = r * r
    st = np.sin(2 * np.pi * r) / r
    basis = np.vstack([x**i for i in range(deg + 1)]).T
    c, e = minimax(basis, st, 1.0 / st, len(r))
    return c, e


def fit_cos(deg):
    r = np.linspace(0.0, 0.125, 4001)
    x = r * r
    ct = np.cos(2 * np.pi * r) - 1.0
    basis = np.vstack([x**i for i in range(1, deg + 1)]).T
    c, e = minimax(basis, ct, 1.0 / np.cos(2 * np.pi * r), len(r))
    return np.concatenate([[1.0], c]), e


if __name__ == "__main__":
    for deg in (6, 7, 8):
        c, e = fit_log(deg)
        print(f"log Q deg {deg}: max rel err {e:.3e}")
        print("   ", ", ".join(hexf(v) + "f" for v in c))
    for deg in (3, 4):
        c, e = fit_sin(deg)
        print(f"sin S deg {deg} (in r^2): max rel err {e:.3e}")
        print("   ", ", ".join(hexf(v) + "f" for v in c))
    for deg in (4, 5):
        c, e = fit_cos(deg)
        print(f"cos C deg {deg} (in r^2): max rel err {e:.3e}")
        print("   ", ", ".join(hexf(v) + "f" for v in c))
