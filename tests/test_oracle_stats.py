"""Statistical parity with the UNMODIFIED reference.

Fixtures (tests/golden/stats_*.csv) hold, per shell, the mean and the
seed-to-seed standard deviation of the reference's heat column
(tiny_mc.c:63) and '# extra' (tiny_mc.c:66), printed by the reference built
from its own sources:

  stats_<phys>_*       the reference as shipped, drawing from glibc rand();
  stats_<phys>_sm_*    the same reference object code with rand()/srand()
                       replaced at link time by SplitMix64
                       (oracle/ref_rand_splitmix.c).

glibc's rand() is additive lagged-Fibonacci (r[i] = r[i-3] + r[i-31]); the
walk draws three to four uniforms per absorption, and the lag structure
biases the shipped reference's far-field heat by about 1% (shells 60-100 at
default physics), which a 2^21-photon run resolves at 5 sigma.  So:

  - against the shipped reference the test asks that our result look like
    one more reference run: z uses the reference's single-run spread;
  - against the SplitMix64-driven reference (same physics, decorrelated
    stream) the test is at full precision: z uses the standard error of the
    fixture's mean;
  - a control test shows the two fixtures differ, i.e. the bias is the
    reference's, not the oracle's.

The shells are positively correlated (one photon deposits in many), which
widens the chi-square's spread; the bound allows ~5 of its standard
deviations under a doubled variance.
"""
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PHYS = {
    "default": dict(mu_a=2.0, mu_s=20.0, shells=101, microns_per_shell=50.0),
    "lowalb": dict(mu_a=10.0, mu_s=10.0, shells=51, microns_per_shell=100.0),
    "highalb": dict(mu_a=0.5, mu_s=50.0, shells=201, microns_per_shell=20.0),
}
FIXTURE = {
    "default": ("stats_default_256x32768.csv", 256, 32768),
    "lowalb": ("stats_lowalb_128x32768.csv", 128, 32768),
    "highalb": ("stats_highalb_128x4096.csv", 128, 4096),
    "default_sm": ("stats_default_sm_8192x8192.csv", 8192, 8192),
    "highalb_sm": ("stats_highalb_sm_8192x1024.csv", 8192, 1024),
}


def load_stats(fixture):
    name, seeds, photons = FIXTURE[fixture]
    mean, sd = [], []
    with open(os.path.join(HERE, "golden", name)) as f:
        for line in f:
            if line.startswith("#"):
                continue
            _, m, s = line.strip().split(",")
            mean.append(float(m))
            sd.append(float(s))
    return np.array(mean), np.array(sd), seeds, photons


def heat_column(heat, photons, mps):
    """tiny_mc.c:60-63 in double: heat / t / (i^2 + i + 1/3), plus extra."""
    S = len(heat)
    t = 4.0 * math.pi * mps**3 * photons / 1e12
    i = np.arange(S - 1, dtype=np.float64)
    col = heat[:-1] / t / (i * i + i + 1.0 / 3.0)
    return np.append(col, heat[-1] / photons)


def z_scores(fixture, heat, photons, per_run):
    mean, sd, seeds, ref_photons = load_stats(fixture)
    mps = PHYS[fixture.replace("_sm", "")]["microns_per_shell"]
    ours = heat_column(heat, photons, mps)
    var = sd**2 / seeds + sd**2 * ref_photons / photons
    if per_run:
        var = var + sd**2
    ok = sd > 0
    return (ours[ok] - mean[ok]) / np.sqrt(var[ok])


def check(z, zmax=5.0):
    dof = z.size
    chi2 = float((z**2).sum())
    assert chi2 < dof + 10 * math.sqrt(2 * dof), (chi2, dof)
    assert np.abs(z).max() < zmax, np.abs(z).max()


CASES = [("default", 1 << 21), ("lowalb", 1 << 21), ("highalb", 1 << 17)]


@pytest.mark.parametrize("sampler", [0, 1])
@pytest.mark.parametrize("variant,photons", CASES)
def test_fp64_oracle_vs_shipped_reference(orc, sampler, variant, photons):
    heat, _, _ = orc.walk64(orc.params(**PHYS[variant]), 20260 + sampler, photons, sampler=sampler)
    check(z_scores(variant, heat, photons, per_run=True), zmax=4.0)


@pytest.mark.parametrize("sampler", [0, 1])
@pytest.mark.parametrize("variant,photons", [("default", 1 << 21), ("highalb", 1 << 17)])
def test_fp64_oracle_vs_reference_physics_decorrelated(orc, sampler, variant, photons):
    heat, _, _ = orc.walk64(orc.params(**PHYS[variant]), 30260 + sampler, photons, sampler=sampler)
    check(z_scores(variant + "_sm", heat, photons, per_run=False))


def test_restatement_passes_its_own_fixture(orc):
    """Calibration: the restatement on the reference's own stream (a seed
    outside the fixture's 1..256) passes the per-run comparison."""
    h, _, _ = orc.restate(orc.params(**PHYS["default"]), 9999, 32768)
    check(z_scores("default", h.astype(np.float64), 32768, per_run=True), zmax=4.0)


@pytest.mark.parametrize("variant", ["default", "highalb"])
def test_shipped_reference_rand_bias_is_resolved(variant):
    """The shipped reference (glibc rand()) and the same code on SplitMix64
    disagree beyond noise in the far field: the bias that makes the
    per-run comparison necessary belongs to the reference's stream."""
    m1, s1, n1, _ = load_stats(variant)
    m2, s2, n2, _ = load_stats(variant + "_sm")
    z = (m2 - m1) / np.sqrt(s1**2 / n1 + s2**2 / n2)
    dof = z.size
    # far beyond the correlated-shell chi-square bound used above ...
    assert (z**2).sum() > dof + 40 * math.sqrt(2 * dof)
    # ... and in the direction of less heat far out (and in the overflow
    # shell) under glibc rand()
    far = slice(dof * 3 // 5, dof - 1)
    assert np.median(z[far]) > 2.0 and z[-1] > 5.0


@pytest.mark.gpu
@pytest.mark.parametrize("sampler", [0, 1])
@pytest.mark.parametrize("variant,photons", [("default", 1 << 24), ("lowalb", 1 << 24), ("highalb", 1 << 22)])
def test_gpu_vs_shipped_reference(tmc, sampler, variant, photons):
    p = tmc.Params(photons=photons, **PHYS[variant])
    tally, _ = tmc.run_raw(p, 777 + sampler, sampler=sampler)
    heat, _ = tmc.tally_to_heat(p, tally)
    check(z_scores(variant, heat, photons, per_run=True), zmax=4.0)


@pytest.mark.gpu
@pytest.mark.parametrize("sampler", [0, 1])
@pytest.mark.parametrize("variant,photons", [("default", 1 << 26), ("highalb", 1 << 23)])
def test_gpu_vs_reference_physics_decorrelated(tmc, sampler, variant, photons):
    p = tmc.Params(photons=photons, **PHYS[variant])
    tally, _ = tmc.run_raw(p, 888 + sampler, sampler=sampler)
    heat, _ = tmc.tally_to_heat(p, tally)
    check(z_scores(variant + "_sm", heat, photons, per_run=False))
