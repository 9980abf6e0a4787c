"""Synthetic parameter sweeps for the batched and served modes.

sweep_sets: the batched-sweep workload — parameter sets with mu_a
log-uniform in [0.05, 10]/cm, mu_s log-uniform in [10, 200]/cm, shells drawn
from {51, 101, 201, 401}, shell width uniform in [10, 100] um, photon budgets
log-uniform in [2^16, 2^22], seed = set index.
"""
import numpy as np

from ._lib import Params

SHELL_CHOICES = (51, 101, 201, 401)


def draw_physics(rng, n):
    mu_a = np.exp(rng.uniform(np.log(0.05), np.log(10.0), n))
    mu_s = np.exp(rng.uniform(np.log(10.0), np.log(200.0), n))
    shells = rng.choice(SHELL_CHOICES, n)
    width = rng.uniform(10.0, 100.0, n)
    return mu_a, mu_s, shells, width


def sweep_sets(n=4096, seed=0, photon_scale=1.0):
    """[(Params, seed)] of the sweep; photon_scale shrinks every budget (>= 64)
    for tests."""
    rng = np.random.default_rng(seed)
    mu_a, mu_s, shells, width = draw_physics(rng, n)
    photons = np.exp(rng.uniform(np.log(2.0**16), np.log(2.0**22), n))
    out = []
    for i in range(n):
        ph = max(64, int(round(photons[i] * photon_scale)))
        out.append((Params(float(np.float32(mu_a[i])), float(np.float32(mu_s[i])), int(shells[i]),
                           float(np.float32(width[i])), ph), i))
    return out
