"""Batched sweep (many parameter sets in one persistent launch): every set's
tallies equal the same set run alone (bit for bit, tiny_mc_gpu_run_ex) and
the oracle's FP32 mirror, under every launch shape, chunking, tally
privatisation and in HIP-graph replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def small_sweep(tmc, n=48, scale=1.0 / 256, seed=5):
    return tmc.sweep_sets(n=n, seed=seed, photon_scale=scale)


def alone(tmc, sets, sampler=0):
    out = []
    for p, seed, *rest in sets:
        t, _ = tmc.run_raw(p, seed, sampler=sampler, first_photon=rest[0] if rest else 0)
        out.append(t)
    return out


@pytest.mark.parametrize("sampler", [0, 1])
def test_batch_equals_sets_alone(tmc, sampler):
    sets = small_sweep(tmc)
    flat, st = tmc.batch_run(sets, tmc.BatchOptions(sampler=sampler))
    parts = tmc.split_tallies(sets, flat)
    ref = alone(tmc, sets, sampler)
    for i, (a, b) in enumerate(zip(parts, ref)):
        np.testing.assert_array_equal(a, b, err_msg=f"set {i}")
    assert st.sets == len(sets) and st.photons == sum(s[0].photons for s in sets)
    # every stepping iteration of a wave accounts for its 64 lanes
    assert (st.interactions + st.idle_lane_steps) % 64 == 0
    assert 0.0 < st.lane_utilisation <= 1.0


@pytest.mark.parametrize("math", [0, 1])
def test_batch_equals_mirror(tmc, orc, hw_tables, math):
    sets = small_sweep(tmc, n=6, scale=1.0 / 1024, seed=9)
    flat, st = tmc.batch_run(sets, tmc.BatchOptions(math=math))
    total = 0
    tables = hw_tables if math == tmc.MATH_HARDWARE else None
    for (p, seed), t in zip(sets, tmc.split_tallies(sets, flat)):
        want, n = orc.mirror32(orc.params(p.mu_a, p.mu_s, p.shells, p.microns_per_shell), seed, p.photons,
                               tables=tables)
        np.testing.assert_array_equal(t, want)
        total += n
    assert st.interactions == total


@pytest.mark.parametrize("opts", [
    dict(block=64), dict(block=1024, copies=1), dict(copies=2), dict(waves_per_simd=2),
    dict(chunk_work=4096), dict(chunk_work=1 << 30), dict(block=512, copies=2, chunk_work=100000),
])
def test_launch_shapes_identical(tmc, opts):
    sets = small_sweep(tmc, n=24, seed=11)
    base, _ = tmc.batch_run(sets)
    got, st = tmc.batch_run(sets, tmc.BatchOptions(**opts))
    np.testing.assert_array_equal(got, base)


def test_first_photon_offsets_and_empty_sets(tmc):
    p = tmc.Params(photons=5000)
    sets = [(p, 3, 0), (tmc.Params(photons=0), 4), (p, 3, 5000), (tmc.Params(photons=1, shells=1), 8),
            (p, 3, (1 << 32) - 2500)]
    flat, st = tmc.batch_run(sets)
    parts = tmc.split_tallies(sets, flat)
    np.testing.assert_array_equal(parts[0], tmc.run_raw(p, 3)[0])
    assert not parts[1].any()
    np.testing.assert_array_equal(parts[2], tmc.run_raw(p, 3, first_photon=5000)[0])
    np.testing.assert_array_equal(parts[3], tmc.run_raw(tmc.Params(photons=1, shells=1), 8)[0])
    np.testing.assert_array_equal(parts[4], tmc.run_raw(p, 3, first_photon=(1 << 32) - 2500)[0])
    flat0, st0 = tmc.batch_run([])
    assert flat0.size == 0 and st0.photons == 0
    # photon index ranges end at 2^63 at most, as in a single run
    far = [(p, 3, (1 << 63) - 5000)]
    np.testing.assert_array_equal(tmc.batch_run(far)[0], tmc.run_raw(p, 3, first_photon=(1 << 63) - 5000)[0])
    with pytest.raises(tmc.TmcError) as e:
        tmc.batch_run([(p, 3, (1 << 63) - 4999)])
    assert e.value.status == -5


def test_graph_replays_identical(tmc):
    """A captured plan replays to the same bits every time, with eager work
    (other launches and synchronisations) between the replays."""
    sets = small_sweep(tmc, n=32, seed=21)
    eager, _ = tmc.batch_run(sets)
    with tmc.BatchPlan(sets, tmc.BatchOptions(use_graph=True)) as plan:
        for r in range(4):
            got, st = plan.run()
            np.testing.assert_array_equal(got, eager, err_msg=f"replay {r}")
            tmc.run_raw(tmc.Params(photons=10000), r)  # eager work in between


def test_autotune_returns_working_options(tmc):
    sets = small_sweep(tmc, n=32, seed=33)
    best = tmc.batch_autotune(sets, sample_divisor=4)
    # each field is either kept automatic (0, as in the default base) or a
    # candidate the tuner timed
    assert best.block in (0, 64, 256, 512)
    assert best.copies in (0, 1, 2)
    assert best.waves_per_simd in (0, 4, 6)
    assert best.chunk_work >= 0
    got, st = tmc.batch_run(sets, best)
    base, _ = tmc.batch_run(sets)
    np.testing.assert_array_equal(got, base)


def test_truncation_in_batch(tmc):
    sets = small_sweep(tmc, n=8, seed=2)
    flat, st = tmc.batch_run(sets, tmc.BatchOptions(max_interactions=10))
    for (p, seed), t in zip(sets, tmc.split_tallies(sets, flat)):
        np.testing.assert_array_equal(t, tmc.run_raw(p, seed, max_interactions=10)[0])
    assert st.truncated > 0


def test_truncation_beyond_flush_interval_in_batch(tmc):
    """Caps below and above the batch kernel's cap-test horizon (2^20 steps,
    the most steps between two slot flushes): equal to the single-run
    kernel (whose own horizon is a few thousand steps).  At albedo 0.99999
    about a tenth of the photons survive the first roulette (~6.9e5
    absorptions) and pass a cap of 2^20 + 5000."""
    sets = [(tmc.Params(mu_a=0.01, mu_s=100.0, shells=401, microns_per_shell=10.0, photons=300), 3),
            (tmc.Params(photons=20000), 4)]
    flat, st = tmc.batch_run(sets, tmc.BatchOptions(max_interactions=20000))
    for (p, seed), t in zip(sets, tmc.split_tallies(sets, flat)):
        np.testing.assert_array_equal(t, tmc.run_raw(p, seed, max_interactions=20000)[0])
    assert st.truncated == 300
    kmax = (1 << 20) + 5000
    sets = [(tmc.Params(mu_a=0.001, mu_s=100.0, shells=101, microns_per_shell=100.0, photons=256), 7)]
    flat, st = tmc.batch_run(sets, tmc.BatchOptions(max_interactions=kmax))
    single, ss = tmc.run_raw(sets[0][0], 7, max_interactions=kmax)
    np.testing.assert_array_equal(flat, single)
    assert st.truncated == ss.truncated and 0 < st.truncated < 256
    assert st.interactions == ss.interactions


def test_max_shells(tmc):
    """LDS slot of 8192 shells (one wave per workgroup) and the single-run
    kernel at the same size."""
    p = tmc.Params(shells=8192, microns_per_shell=1.0, photons=20000)
    flat, st = tmc.batch_run([(p, 1)])
    assert st.block == 64
    single, _ = tmc.run_raw(p, 1)
    np.testing.assert_array_equal(flat, single)


def test_lds_class_passes(tmc):
    """Sets whose shell counts allow different resident waves run as separate
    launches (one per LDS size class carrying at least a tenth of the work),
    and every set still equals the same set run alone."""
    small = [(tmc.Params(shells=51, microns_per_shell=100.0, photons=40000), s) for s in range(4)]
    big = [(tmc.Params(shells=401, microns_per_shell=12.5, photons=40000), 10 + s) for s in range(4)]
    huge = [(tmc.Params(shells=2001, microns_per_shell=2.5, photons=60000), 20)]
    for sets, passes in ((small + big, 2), (small + big[:1] + [(tmc.Params(shells=401, photons=500), 30)], 2),
                         (small + [(tmc.Params(shells=401, photons=2000), 31)], 1), (small + big + huge, 3)):
        flat, st = tmc.batch_run(sets)
        assert st.passes == passes, (passes, st)
        for i, (a, b) in enumerate(zip(tmc.split_tallies(sets, flat), alone(tmc, sets))):
            np.testing.assert_array_equal(a, b, err_msg=f"set {i} of a {passes}-pass batch")
        assert st.photons == sum(p.photons for p, _ in sets)
        assert (st.interactions + st.idle_lane_steps) % 64 == 0
