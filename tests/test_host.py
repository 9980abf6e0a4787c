"""Host-side logic of the library and the package (CPU only, no device work).

- tiny_mc_gpu_expected_interactions, the estimate the batched and served
  modes schedule with, equals the FP32 walk's mean absorptions per photon
  (the oracle's bit-exact mirror of the kernel) within 4 standard errors,
  across albedos from 0.5 to 0.9999;
- the sweep and serving workloads are the stated distributions, and the
  sweep is pinned (its total photon budget at n=4096, seed 0);
- the batch tally layout (tmc_batch_tally_words, split_tallies) is the
  concatenation of 4 * shells words per set;
- tmc_serve_open rejects invalid options before any device work.
"""
import ctypes

import numpy as np
import pytest

PHYS = [
    dict(mu_a=2.0, mu_s=20.0, shells=101, microns_per_shell=50.0),  # default
    dict(mu_a=10.0, mu_s=10.0, shells=51, microns_per_shell=100.0),  # albedo 0.5
    dict(mu_a=0.5, mu_s=50.0, shells=201, microns_per_shell=20.0),
    dict(mu_a=0.01, mu_s=100.0, shells=401, microns_per_shell=10.0),  # precision stress
]


@pytest.mark.parametrize("phys,photons", [(PHYS[0], 1 << 16), (PHYS[1], 1 << 16), (PHYS[2], 1 << 14),
                                          (PHYS[3], 1 << 8)])
def test_expected_interactions_match_the_fp32_walk(tmc, orc, phys, photons):
    L = tmc.expected_interactions(tmc.Params(**phys))
    per = np.array([orc.mirror32(orc.params(**phys), 7, 1, first=i, nthreads=1)[1] for i in range(64)]
                   if photons <= 256 else [])
    _, n = orc.mirror32(orc.params(**phys), 7, photons)
    mean = n / photons
    # per-photon spread: one roulette cycle of k2 absorptions per survival,
    # survival probability 0.1 -> sd of the count ~ k2 * sqrt(0.1) / 0.9
    sd = (per.std() if per.size else 0.0) or L * 0.15
    assert abs(mean - L) < 4.0 * sd / np.sqrt(photons), (mean, L)


def test_expected_interactions_closed_forms(tmc):
    # default physics: first cutoff after 73 absorptions, then cycles of
    # 24 (25 every few cycles); mean = k1 + k2 p / (1 - p) with p the FP32
    # roulette survival probability
    L = tmc.expected_interactions(tmc.Params())
    assert abs(L - 75.6667) < 2e-3
    # precision stress: FP32 cutoff at absorption 69,063, then every 23,021
    L = tmc.expected_interactions(tmc.Params(**PHYS[3]))
    assert abs(L - (69063 + 23021 * 0.1 / 0.9)) < 2.0
    # invalid parameters give 0, no exception
    assert tmc.expected_interactions(tmc.Params(mu_a=0.0)) == 0.0


def test_sweep_sets_distribution_and_pin(tmc):
    sets = tmc.sweep_sets(n=4096, seed=0)
    assert len(sets) == 4096 and [s for _, s in sets] == list(range(4096))
    p = sets
    mu_a = np.array([q.mu_a for q, _ in p])
    mu_s = np.array([q.mu_s for q, _ in p])
    width = np.array([q.microns_per_shell for q, _ in p])
    photons = np.array([q.photons for q, _ in p])
    assert set(q.shells for q, _ in p) == {51, 101, 201, 401}
    assert mu_a.min() >= 0.05 * (1 - 1e-6) and mu_a.max() <= 10.0 * (1 + 1e-6)
    assert mu_s.min() >= 10.0 * (1 - 1e-6) and mu_s.max() <= 200.0 * (1 + 1e-6)
    assert width.min() >= 10.0 and width.max() <= 100.0
    assert photons.min() >= 2**16 and photons.max() <= 2**22
    # log-uniform: the mean of log is the midpoint
    assert abs(np.log(mu_a).mean() - 0.5 * (np.log(0.05) + np.log(10.0))) < 0.1
    # the workload the bench reports
    assert int(photons.sum()) == 4055528768
    # scaled copies keep at least 64 photons per set
    small = tmc.sweep_sets(n=16, seed=1, photon_scale=1e-6)
    assert min(q.photons for q, _ in small) == 64


def test_serving_requests_distribution(tmc):
    reqs = tmc.serve_load.draw_requests(20000, seed=3)
    photons = np.array([q.photons for q, _ in reqs])
    assert photons.min() >= 2**14 and photons.max() <= 2**24
    # Pareto(1.5, 2^14): median 2^14 * 2^(1/1.5) = 26,008
    assert abs(np.median(photons) / 26008 - 1.0) < 0.05
    assert len({s for _, s in reqs}) == len(reqs)
    assert tmc.serve_load.mean_request_work(reqs[:100]) > 0
    # the sweep's physics: mu_a log-uniform on [0.05, 10], mu_s on [10, 200], four shell counts
    mu_a = np.array([q.mu_a for q, _ in reqs])
    assert mu_a.min() >= 0.05 * (1 - 1e-6) and mu_a.max() <= 10.0 * (1 + 1e-6)
    assert abs(np.median(np.log(mu_a)) - 0.5 * (np.log(0.05) + np.log(10.0))) < 0.05
    assert {q.shells for q, _ in reqs} == set(tmc.sweep.SHELL_CHOICES)
    # request i depends on (seed, i) only: a shorter list is a prefix
    assert reqs[:1234] == tmc.serve_load.draw_requests(1234, seed=3)


def test_batch_tally_layout(tmc):
    from tiny_mc_amd._lib import _bind_batch, _Set

    lib = tmc.load()
    _bind_batch(lib)
    sets = tmc.sweep_sets(n=5, seed=2)
    arr = (_Set * len(sets))()
    for i, (q, s) in enumerate(sets):
        arr[i] = _Set(q._c(), s, 0)
    want = sum(4 * q.shells for q, _ in sets)
    assert lib.tmc_batch_tally_words(arr, len(sets)) == want
    flat = np.arange(want, dtype=np.uint64)
    parts = tmc.split_tallies(sets, flat)
    assert [len(x) for x in parts] == [4 * q.shells for q, _ in sets]
    assert np.array_equal(np.concatenate(parts), flat)
    assert lib.tmc_batch_tally_words(None, 3) == 0


@pytest.mark.parametrize("kw", [dict(sampler=2), dict(math=2), dict(max_shells=8193), dict(block=96),
                                dict(block=1088), dict(ring_chunks=1000), dict(max_requests=65537),
                                dict(max_photons_per_lane=(1 << 24) + 1)])
def test_serve_open_rejects_invalid_options(tmc, kw):
    """Invalid serving options fail with TMC_E_INVALID before the server
    touches a device, so on any host.  The per-lane chunk cap keeps a chunk's
    photon count (at most 64 * max_photons_per_lane) within its 32-bit
    field."""
    with pytest.raises(tmc.TmcError) as e:
        tmc.Server(**kw)
    assert e.value.status == -1
