"""Parity of the HIP walk kernel with the oracle (GPU tests).

Bit-exact: the kernel's 128-bit integer tallies equal the FP32 mirror's
(oracle.c orc_mirror32) on the same Philox stream, for both direction
samplers, both math modes and across physics, shell counts, photon-index
offsets and launch shapes.  In the hardware-math mode (the default) the
mirror takes the device's step length and sin/cos for each of the 2^24
uniforms the walk can draw (their accuracy against float64 is checked here
on every one of them) and restates every other operation.  Tolerance: at
2^24 photons every shell's heat is within max(3 sigma from heat2, 1e-5
relative) of the FP64 oracle (orc_walk64)."""
import numpy as np
import pytest

import oracle_lib

pytestmark = pytest.mark.gpu

DEFAULT = dict(mu_a=2.0, mu_s=20.0, shells=101, microns_per_shell=50.0)
LOWALB = dict(mu_a=10.0, mu_s=10.0, shells=51, microns_per_shell=100.0)
HIGHALB = dict(mu_a=0.5, mu_s=50.0, shells=201, microns_per_shell=20.0)


def _params(tmc, phys, photons):
    return tmc.Params(photons=photons, **phys)


def _mirror(orc, phys, seed, photons, **kw):
    return orc.mirror32(orc.params(**phys), seed, photons, **kw)


@pytest.mark.parametrize("math", [0, 1])
@pytest.mark.parametrize("sampler", [0, 1])
@pytest.mark.parametrize("phys,seed,photons,first", [
    (DEFAULT, 12345, 1 << 16, 0),
    (DEFAULT, (1 << 40) + 5, 20000, (1 << 32) - 7000),   # photon index crosses 2^32
    (LOWALB, 3, 1 << 16, 17),
    (HIGHALB, 11, 4096, 0),
    (dict(DEFAULT, shells=1), 9, 5000, 0),
    (dict(DEFAULT, shells=2), 9, 5000, 0),
    (dict(DEFAULT, shells=401, microns_per_shell=10.0), 77, 10000, 0),
])
def test_tallies_bit_exact_vs_mirror(tmc, orc, hw_tables, math, sampler, phys, seed, photons, first):
    tally, st = tmc.run_raw(_params(tmc, phys, photons), seed, sampler=sampler, first_photon=first, math=math)
    assert st.math == math
    tables = hw_tables if math == tmc.MATH_HARDWARE else None
    want, n = _mirror(orc, phys, seed, photons, first=first, sampler=sampler, tables=tables)
    assert st.interactions == n
    assert st.photons == photons and st.truncated == 0
    np.testing.assert_array_equal(tally, want)


@pytest.mark.parametrize("math", [0, 1])
@pytest.mark.parametrize("sampler", [0, 1])
@pytest.mark.parametrize("phys,photons,table", [
    (DEFAULT, 1 << 20, 0),          # the automatic table (649 entries at default physics)
    (DEFAULT, 1 << 16, 80),         # forced, ends 7 absorptions into the second cycle: survivors fall back
    (LOWALB, 1 << 18, 12),          # albedo 0.5: k1 = 10, the table ends inside the second cycle
    (HIGHALB, 1 << 14, 1024),       # k1 = 695: survivors of two roulettes run past the table
])
def test_weight_table_bit_exact(tmc, orc, hw_tables, math, sampler, phys, photons, table):
    """The walk that reads its weight images from the LDS table equals the
    mirror (which multiplies) bit for bit, also for photons that outlive
    the table and continue with the arithmetic from its last weight."""
    tally, st = tmc.run_raw(_params(tmc, phys, photons), 606, sampler=sampler, math=math, weight_table=table)
    assert st.weight_table == (table or 649)
    tables = hw_tables if math == tmc.MATH_HARDWARE else None
    want, n = _mirror(orc, phys, 606, photons, sampler=sampler, tables=tables)
    assert st.interactions == n
    np.testing.assert_array_equal(tally, want)


def test_weight_table_choice(tmc):
    """Automatic table: on for runs of 2^20 photons or more (either
    sampler) whose table spans nine roulette cycles within 1024 entries, off
    otherwise; never for the other precision modes."""
    big = 1 << 20
    assert tmc.run_raw(_params(tmc, DEFAULT, big), 1)[1].weight_table == 73 + 24 * 24
    assert tmc.run_raw(_params(tmc, DEFAULT, big - 64), 1)[1].weight_table == 0
    assert tmc.run_raw(_params(tmc, HIGHALB, big), 1, max_interactions=64)[1].weight_table == 0
    assert tmc.run_raw(_params(tmc, DEFAULT, big), 1, weight_table=tmc.WEIGHT_TABLE_OFF)[1].weight_table == 0
    assert tmc.run_raw(_params(tmc, DEFAULT, big), 1, sampler=tmc.SAMPLER_REJECTION)[1].weight_table == 649
    assert tmc.run_raw(_params(tmc, DEFAULT, 4096), 1, precision=tmc.PRECISION_FP64, weight_table=64)[1].weight_table == 0


@pytest.mark.parametrize("block,grid,chunk", [(64, 1, 64), (256, 7, 128), (1024, 0, 4096), (512, 3000, 64)])
def test_launch_shape_independence(tmc, orc, hw_tables, block, grid, chunk):
    p = _params(tmc, DEFAULT, 50000)
    want, n = _mirror(orc, DEFAULT, 4242, 50000, tables=hw_tables)
    tally, st = tmc.run_raw(p, 4242, block=block, grid=grid, chunk=chunk)
    np.testing.assert_array_equal(tally, want)
    assert st.interactions == n


def test_truncation_matches_mirror(tmc, orc, hw_tables):
    p = _params(tmc, DEFAULT, 30000)
    for kmax in (1, 2, 50):
        tally, st = tmc.run_raw(p, 5, max_interactions=kmax)
        want, n = _mirror(orc, DEFAULT, 5, 30000, max_k=kmax, tables=hw_tables)
        np.testing.assert_array_equal(tally, want)
        assert st.interactions == n
        if kmax == 1:
            assert st.truncated == 30000 and n == 30000


def test_truncation_beyond_flush_interval(tmc, orc, hw_tables):
    """Caps larger than the steps between two tally flushes: the walk checks
    the cap only from the flush before some photon can reach it, and the
    truncated walk still equals the mirror's (at albedo 0.9999 every photon
    reaches 20000 absorptions; at 69070 the roulette at absorption 69063
    kills most of them first)."""
    phys = dict(mu_a=0.01, mu_s=100.0, shells=401, microns_per_shell=10.0)
    for kmax, photons in ((20000, 512), (69070, 256)):
        p = _params(tmc, phys, photons)
        tally, st = tmc.run_raw(p, 8, max_interactions=kmax)
        want, n = _mirror(orc, phys, 8, photons, max_k=kmax, tables=hw_tables)
        np.testing.assert_array_equal(tally, want)
        assert st.interactions == n
        if kmax == 20000:
            assert st.truncated == photons and n == photons * kmax
        else:
            assert 0 < st.truncated < photons


def test_run_twice_identical_and_resume_adds(tmc):
    p = _params(tmc, DEFAULT, 300000)
    a, _ = tmc.run_raw(p, 99)
    b, _ = tmc.run_raw(p, 99)
    np.testing.assert_array_equal(a, b)
    # [0, N) == [0, A) + [A, N) as 128-bit integers
    A = 123457
    lo, _ = tmc.run_raw(_params(tmc, DEFAULT, A), 99)
    hi, _ = tmc.run_raw(_params(tmc, DEFAULT, 300000 - A), 99, first_photon=A)
    s = lo.astype(object) + hi.astype(object)
    s = np.array([int(v) for v in s], dtype=object)
    tot = [(int(s[2 * i]) + (int(s[2 * i + 1]) << 64)) for i in range(2 * 101)]
    ref = [(int(a[2 * i]) + (int(a[2 * i + 1]) << 64)) for i in range(2 * 101)]
    assert tot == ref


def test_zero_photons_and_bad_params(tmc):
    t, st = tmc.run_raw(_params(tmc, DEFAULT, 0), 1)
    assert not t.any() and st.interactions == 0
    for bad in (dict(DEFAULT, mu_a=0.0), dict(DEFAULT, shells=0), dict(DEFAULT, microns_per_shell=-1.0),
                dict(DEFAULT, mu_s=-1.0), dict(DEFAULT, shells=tmc.MAX_SHELLS + 1),
                dict(DEFAULT, mu_a=1e-9, mu_s=100.0)):
        with pytest.raises(tmc.TmcError) as e:
            tmc.run_raw(_params(tmc, bad, 10), 1)
        assert e.value.status == -1
    # photon index ranges end at 2^63 at most
    for first in ((1 << 64) - 5, (1 << 63) - 5):
        with pytest.raises(tmc.TmcError) as e:
            tmc.run_raw(_params(tmc, DEFAULT, 10), 1, first_photon=first)
        assert e.value.status == -5
    t, st = tmc.run_raw(_params(tmc, DEFAULT, 10), 1, first_photon=(1 << 63) - 10)
    assert st.photons == 10 and t.any()


def test_heat_api_matches_tally(tmc):
    p = _params(tmc, DEFAULT, 1 << 16)
    heat, heat2 = tmc.run(p, 31)
    tally, _ = tmc.run_raw(p, 31)
    h, h2 = oracle_lib.tally_heat(tally, 101, 2.0, 20.0)
    np.testing.assert_allclose(heat, h, rtol=1e-15, atol=0)
    np.testing.assert_allclose(heat2, h2, rtol=1e-15, atol=0)


def test_device_math_bit_exact_exhaustive(tmc, orc):
    """log and sin/cos(2 pi a) on every uniform the walk can draw."""
    x = (np.arange(1, (1 << 24) + 1, dtype=np.float64) * 2.0 ** -24).astype(np.float32)
    lg, sn, cs = tmc.eval_math(x)
    ol, osn, ocs = orc.eval_math32(x)
    np.testing.assert_array_equal(lg.view(np.uint32), ol.view(np.uint32))
    np.testing.assert_array_equal(sn.view(np.uint32), osn.view(np.uint32))
    np.testing.assert_array_equal(cs.view(np.uint32), ocs.view(np.uint32))


@pytest.mark.parametrize("sampler", [0, 1])
def test_vs_fp64_oracle_2e24(tmc, orc, sampler):
    """Every shell within max(3 sigma from heat2, 1e-5 relative) of the FP64
    oracle on the same stream, photons [0, 2^24)."""
    N = 1 << 24
    p = _params(tmc, DEFAULT, N)
    tally, st = tmc.run_raw(p, 2024, sampler=sampler)
    heat, heat2 = tmc.tally_to_heat(p, tally)
    oh, oh2, on = orc.walk64(orc.params(**DEFAULT), 2024, N, sampler=sampler)
    sigma = np.sqrt(np.maximum(oh2 - oh * oh / N, 0.0))
    tol = np.maximum(3.0 * sigma, 1e-5 * np.abs(oh))
    bad = np.nonzero(np.abs(heat - oh) > tol)[0]
    assert bad.size == 0, (bad, heat[bad], oh[bad], tol[bad])
    # same roulette decisions: the FP32 and FP64 weight cadences agree at
    # default physics, so the absorption counts are identical
    assert st.interactions == on
    # and the aggregate is rounding-level close
    assert abs(heat.sum() / oh.sum() - 1.0) < 1e-6


def test_short_sqrt_exhaustive(tmc):
    """The walk's short correctly rounded sqrt equals the IEEE sqrt on every
    float of its domain: +0 and [2^-96, FLT_MAX] (all 1.8e9 patterns)."""
    lo = int(np.float32(2.0 ** -96).view(np.uint32))
    hi = int(np.float32(np.finfo(np.float32).max).view(np.uint32))
    assert tmc.check_sqrt(0, 1) == (0, 0)
    assert tmc.check_sqrt(lo, hi - lo + 1) == (0, 0)


@pytest.mark.parametrize("phys", [dict(), dict(mu_a=10.0, mu_s=10.0, shells=51, microns_per_shell=100.0),
                                  dict(mu_a=0.5, mu_s=50.0, shells=201, microns_per_shell=20.0),
                                  dict(mu_a=0.01, mu_s=100.0, shells=401, microns_per_shell=10.0),
                                  dict(shells=8192, microns_per_shell=1.0)])
def test_shell_index_claim_exhaustive(tmc, phys):
    """Where the walk takes a deposit's shell from the hardware square root,
    the shell and the overflow side are those of the correctly rounded
    radius, on every non-negative float squared radius (+inf included);
    the other radii are recomputed with the IEEE square root."""
    inf = int(np.float32(np.inf).view(np.uint32))
    bad, first, unsure = tmc.check_shell_index(tmc.Params(**phys), 0, inf + 1)
    assert (bad, first) == (0, 0)
    # the fallback covers the tiny radii (about half of all float patterns
    # lie below 2^-1) and the neighbourhoods of the shell boundaries
    assert 0 < unsure < inf


def test_roulette_division_exhaustive(tmc):
    """The roulette's weight update w / 0.1f as fma(w, 10, w * C_lo) equals
    the IEEE quotient on every float of magnitude from 2^-100 to the largest
    finite one, and the guarded form the walk calls equals it on every
    finite float (zeros and subnormals included; lanes without a photon
    carry negative weights through the step)."""
    lo = int(np.float32(2.0 ** -100).view(np.uint32))
    inf = int(np.float32(np.inf).view(np.uint32))
    sign = 1 << 31
    for base in (0, sign):
        assert tmc.check_short_op(tmc.CHECK_DIV_TENTH_SHORT, base + lo, inf - lo) == (0, 0)
        assert tmc.check_short_op(tmc.CHECK_DIV_TENTH, base, inf) == (0, 0)
    # the short form alone is not exact everywhere: the guard is needed
    n, first = tmc.check_short_op(tmc.CHECK_DIV_TENTH_SHORT, 0, lo)
    assert n > 0 and first > 0


def test_hardware_math_accuracy_exhaustive(tmc, orc, hw_tables):
    """Hardware-math step length -ln(u) within 2 ulp, sin/cos(2 pi u) within
    2e-7 and the analytic sampler's sin(theta) = sqrt(1 - (2u - 1)^2)
    within 2 ulp of float64 on every uniform the walk can draw; the own-math
    evaluation equals the host polynomials and the correctly rounded
    sin(theta) bit for bit."""
    t, sn, cs, sp = hw_tables
    m = np.arange(1, (1 << 24) + 1, dtype=np.float64)
    u = m * 2.0 ** -24

    def ulps(got, ref):
        ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
        ulp[ulp == 0] = np.spacing(np.float32(0))
        return np.abs(got - ref) / ulp

    assert np.max(ulps(t, -np.log(u))) <= 2.0
    assert np.max(np.abs(sn - np.sin(2 * np.pi * u))) <= 2e-7
    assert np.max(np.abs(cs - np.cos(2 * np.pi * u))) <= 2e-7
    mu = 2.0 * u - 1.0  # exact, as is 1 - mu^2
    assert np.max(ulps(sp, np.sqrt(1.0 - mu * mu))) <= 2.0
    # the own polynomials as the walk evaluates them, against the host's
    idx = np.arange(0, 1 << 24, 97)
    uf = u[idx].astype(np.float32)
    ot, osn, ocs, osp = tmc.eval_step_math(uf, tmc.MATH_OWN)
    lg, hs, hc = orc.eval_math32(uf)
    np.testing.assert_array_equal(ot.view(np.uint32), (-lg).view(np.uint32))
    np.testing.assert_array_equal(osn.view(np.uint32), hs.view(np.uint32))
    np.testing.assert_array_equal(ocs.view(np.uint32), hc.view(np.uint32))
    want_sp = np.sqrt((1.0 - mu[idx] * mu[idx]).astype(np.float32))
    np.testing.assert_array_equal(osp.view(np.uint32), want_sp.view(np.uint32))
