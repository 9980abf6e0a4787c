"""Parity beyond default physics, at the max(3 sigma, 1e-5 relative) bar, and
size-independent properties of the full-size run (GPU tests).

- FP32 GPU walk vs FP64 oracle (orc_walk64, same Philox stream): every
  shell's heat within max(3 sigma, 1e-5 relative), for both direction
  samplers, at albedo 0.5 (2^24 photons), at mu_a = 0.5, mu_s = 50 over
  201 x 20 um shells (about 720 absorptions per photon, 2^20 photons) and
  at albedo 0.05, where roulette survivors re-enter below the 0.001 cutoff
  (photon.c:45-49: w / 0.1 of a weight below 0.0001) and their fixed-point
  images are truncated.  sigma is the larger of the reference's estimate
  from heat2 (tiny_mc.c:63, which treats every deposit as independent: with
  hundreds of absorptions per photon a photon deposits in a shell many
  times, and its radicand goes negative) and the standard error of the
  shell sum from 16 independent photon batches of the oracle.  On the same
  stream the two differ only by rounding, so every shell with at least
  1e-3 of the largest shell's heat must also agree within 2e-4 relative
  (deposits that land on the other side of a shell boundary under binary32
  positions: about 1e-5 relative at 720 absorptions per photon).
- Bit-exact against the FP32 mirror in that truncated regime (albedo 0.05
  and 0.09).
- The full-size workload (2^30 photons, default physics, one launch): the
  mean absorptions per photon, the mean absorbed energy and the mean sum
  of squared deposits each within 5 standard errors of their exact
  expectations, computed here from the binary32 weight sequence of
  photon.c:30-32,45-49 (the cadence of cycles and the survival probability
  1677721 / 2^24 of the roulette draw); the heat sums' high words are zero
  and the heat2 sums' carries into theirs are counted exactly (a lost carry
  would be off by 2^64 units, ~10^7 standard errors).
"""
import math

import numpy as np
import pytest

import oracle_lib

pytestmark = pytest.mark.gpu

CASES = {
    "albedo0.5": (dict(mu_a=10.0, mu_s=10.0, shells=51, microns_per_shell=100.0), 1 << 24),
    "mua0.5_mus50_201": (dict(mu_a=0.5, mu_s=50.0, shells=201, microns_per_shell=20.0), 1 << 20),
    "albedo0.05": (dict(mu_a=19.0, mu_s=1.0, shells=51, microns_per_shell=100.0), 1 << 22),
}


@pytest.mark.parametrize("sampler", [0, 1])
@pytest.mark.parametrize("case", sorted(CASES))
def test_vs_fp64_oracle_wide(tmc, orc, case, sampler):
    phys, N = CASES[case]
    print(f"fp32 GPU vs fp64 oracle: {case} N={N} sampler={sampler}", flush=True)
    p = tmc.Params(photons=N, **phys)
    tally, st = tmc.run_raw(p, 777, sampler=sampler)
    heat, _ = tmc.tally_to_heat(p, tally)
    B = 16
    parts = [orc.walk64(orc.params(**phys), 777, N // B, first=b * (N // B), sampler=sampler) for b in range(B)]
    batches = np.array([h for h, _, _ in parts])
    oh = batches.sum(axis=0)
    oh2 = np.sum([h2 for _, h2, _ in parts], axis=0)
    on = sum(n for _, _, n in parts)
    sigma_heat2 = np.sqrt(np.maximum(oh2 - oh * oh / N, 0.0))
    sigma_batch = np.sqrt(B) * batches.std(axis=0, ddof=1)
    sigma = np.maximum(sigma_heat2, sigma_batch)
    tol = np.maximum(3.0 * sigma, 1e-5 * np.abs(oh))
    bad = np.nonzero(np.abs(heat - oh) > tol)[0]
    assert bad.size == 0, (bad, heat[bad], oh[bad], tol[bad])
    big = oh >= 1e-3 * oh.max()
    rel = np.abs(heat[big] / oh[big] - 1.0)
    print(f"  max same-stream relative difference {rel.max():.3g}", flush=True)
    assert rel.max() < 2e-4
    assert abs(st.interactions / on - 1.0) < 1e-4
    assert abs(heat.sum() / oh.sum() - 1.0) < 1e-5


@pytest.mark.parametrize("sampler", [0, 1])
@pytest.mark.parametrize("mu_a,mu_s", [(19.0, 1.0), (9.1, 0.9)])
def test_truncated_images_bit_exact(tmc, orc, hw_tables, mu_a, mu_s, sampler):
    """Albedo 0.05 and 0.09: deposits below 2^-10 (truncated fixed-point
    images) are reproduced by the mirror bit for bit."""
    phys = dict(mu_a=mu_a, mu_s=mu_s, shells=51, microns_per_shell=100.0)
    N = 1 << 18
    tally, st = tmc.run_raw(tmc.Params(photons=N, **phys), 99, sampler=sampler)
    want, n = orc.mirror32(orc.params(**phys), 99, N, sampler=sampler, tables=hw_tables)
    assert st.interactions == n
    np.testing.assert_array_equal(tally, want)


def u128(lo, hi):
    return [int(a) + (int(b) << 64) for a, b in zip(lo, hi)]


def test_full_size_properties_2e30(tmc):
    """2^30 photons at default physics in one launch (the measured
    workload)."""
    N = 1 << 30
    p = tmc.Params(photons=N)
    print(f"full-size run: N={N}", flush=True)
    tally, st = tmc.run_raw(p, 20241203)
    (mk, vk), (me, ve), (mq, vq) = oracle_lib.cadence_moments(2.0, 20.0)
    S = 101
    assert st.photons == N and st.truncated == 0
    # absorptions per photon
    assert abs(st.interactions / N - mk) < 5 * math.sqrt(vk / N), (st.interactions / N, mk)
    # energy: heat sums below 2^64 (max shell ~0.13 * 2^64 at this size)
    assert not tally[1:2 * S:2].any()
    heat_total = sum(u128(tally[0:2 * S:2], tally[1:2 * S:2]))
    energy = (1.0 - 20.0 / 22.0) * heat_total / 2.0**33 / N
    assert abs(energy - me) < 5 * math.sqrt(ve / N), (energy, me)
    assert abs(me - 1.0) < 1e-6  # binary32 weight decay shifts the mean by ~3.6e-7
    # squared deposits: the busiest shells' sums pass 2^64 many times over
    lo2, hi2 = tally[2 * S::2], tally[2 * S + 1::2]
    assert hi2.max() > 0
    sq_total = sum(u128(lo2, hi2))
    assert abs(sq_total / N - mq) < 5 * math.sqrt(vq / N), (sq_total / N, mq)
    # the same totals through the library's double conversion
    heat, heat2 = tmc.tally_to_heat(p, tally)
    assert abs(heat.sum() / N - energy) < 1e-12
    assert abs(heat2.sum() / ((1.0 - 20.0 / 22.0) ** 2 * sq_total / 2.0**43) - 1.0) < 1e-12
