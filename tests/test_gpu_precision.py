"""Precision variants of the walk (TMC_PRECISION_*), GPU tests.

Stress physics: mu_a = 0.01/cm, mu_s = 100/cm, 401 shells x 10 um (albedo
0.9999, about 7e4 absorptions per photon), where binary32 arithmetic
drifts the most.

- FP64: the binary64 GPU walk agrees with the FP64 oracle (orc_walk64) on
  the same stream: identical absorption counts, and per-shell sums equal up
  to the rounding of the fixed-point images and the rare deposit that a
  last-place difference of log/sincos moves across a shell boundary.
- FP32 with the weight from exact counts (with and without compensated
  positions): each photon's roulette cadence is FP64's (cutoff at absorption
  69082, next at 92109), so per-photon absorption counts equal the FP64
  oracle's, where the reference's FP32 arithmetic (cutoff at 69063, equal to
  the bit-exact mirror) matches none.
- At default physics every variant is within the FP32 tolerance
  max(3 sigma, 1e-5 relative) of the FP64 oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STRESS = dict(mu_a=0.01, mu_s=100.0, shells=401, microns_per_shell=10.0)
DEFAULT = dict(mu_a=2.0, mu_s=20.0, shells=101, microns_per_shell=50.0)


def _one_minus_albedo(phys):
    mu_a, mu_s = float(np.float32(phys["mu_a"])), float(np.float32(phys["mu_s"]))
    return 1.0 - mu_s / (mu_s + mu_a)


@pytest.mark.parametrize("sampler", [0, 1])
@pytest.mark.parametrize("phys,photons,seed", [(STRESS, 1 << 11, 5), (DEFAULT, 1 << 16, 12)])
def test_fp64_vs_oracle(tmc, orc, sampler, phys, photons, seed):
    p = tmc.Params(photons=photons, **phys)
    tally, st = tmc.run_raw(p, seed, sampler=sampler, precision=tmc.PRECISION_FP64)
    assert st.precision == tmc.PRECISION_FP64
    heat, heat2 = tmc.tally_to_heat(p, tally)
    oh, oh2, on = orc.walk64(orc.params(**phys), seed, photons, sampler=sampler)
    # the same weights and roulette draws: the same absorption counts
    assert st.interactions == on
    oma = _one_minus_albedo(phys)
    # rounding of the images (relative) plus four deposits (weight <= 1)
    # moved across a shell boundary
    tol = 1e-8 * np.abs(oh) + 4.0 * oma
    tol2 = 1e-8 * np.abs(oh2) + 4.0 * oma * oma
    bad = np.nonzero(np.abs(heat - oh) > tol)[0]
    assert bad.size == 0, (bad, heat[bad], oh[bad])
    bad2 = np.nonzero(np.abs(heat2 - oh2) > tol2)[0]
    assert bad2.size == 0, (bad2, heat2[bad2], oh2[bad2])
    assert abs(heat.sum() / oh.sum() - 1.0) < 1e-9


def _per_photon(tmc, precision, n, seed):
    p = tmc.Params(photons=1, **STRESS)
    return [tmc.run_raw(p, seed, first_photon=i, precision=precision)[1].interactions for i in range(n)]


def test_weight_cadence_matches_fp64(tmc, orc):
    n, seed = 16, 5
    op = orc.params(**STRESS)
    want = [orc.walk64(op, seed, 1, first=i, nthreads=1)[2] for i in range(n)]
    assert min(want) == 69082  # the FP64 cutoff absorption
    assert _per_photon(tmc, tmc.PRECISION_FP64, n, seed) == want
    for prec in (tmc.PRECISION_FP32_WEIGHT, tmc.PRECISION_FP32_COMPENSATED):
        got = _per_photon(tmc, prec, n, seed)
        assert sum(g == w for g, w in zip(got, want)) >= n - 1, (prec, got, want)
    plain = _per_photon(tmc, tmc.PRECISION_FP32, n, seed)
    assert plain == [orc.mirror32(op, seed, 1, first=i, nthreads=1)[1] for i in range(n)]
    assert min(plain) == 69063
    assert not any(g == w for g, w in zip(plain, want))


def test_mitigated_energy_tracks_fp64_at_stress(tmc):
    """Same roulette decisions (but for the rare photon whose weight lies
    within float rounding of the cutoff): the absorbed energy of the
    mitigated FP32 walks is within 5e-5 of FP64's, where the reference's
    FP32 weight decay biases it by about -2.7e-4."""
    N, seed = 1 << 12, 21
    p = tmc.Params(photons=N, **STRESS)
    ref = tmc.tally_to_heat(p, tmc.run_raw(p, seed, precision=tmc.PRECISION_FP64)[0])[0].sum()
    for prec in (tmc.PRECISION_FP32_WEIGHT, tmc.PRECISION_FP32_COMPENSATED):
        e = tmc.tally_to_heat(p, tmc.run_raw(p, seed, precision=prec)[0])[0].sum()
        assert abs(e / ref - 1.0) < 5e-5, (prec, e / ref - 1.0)


@pytest.mark.parametrize("precision", [1, 2, 3])
def test_default_physics_within_fp32_tolerance(tmc, orc, precision):
    N, seed = 1 << 20, 2024
    p = tmc.Params(photons=N, **DEFAULT)
    tally, st = tmc.run_raw(p, seed, precision=precision)
    heat, _ = tmc.tally_to_heat(p, tally)
    oh, oh2, on = orc.walk64(orc.params(**DEFAULT), seed, N)
    sigma = np.sqrt(np.maximum(oh2 - oh * oh / N, 0.0))
    tol = np.maximum(3.0 * sigma, 1e-5 * np.abs(oh))
    bad = np.nonzero(np.abs(heat - oh) > tol)[0]
    assert bad.size == 0, (bad, heat[bad], oh[bad], tol[bad])
    assert st.interactions == on


def test_invalid_precision_rejected(tmc):
    with pytest.raises(tmc.TmcError) as e:
        tmc.run_raw(tmc.Params(photons=64), 1, precision=4)
    assert e.value.status == -1
