"""Closed-form checks that pin the FP64 oracle (orc_walk64) and the FP32
mirror (orc_mirror32) on the Philox stream, where the reference's own
stream cannot be used: weight cadence, absorption count, energy balance,
first-step shell law, direction moments, agreement of the two precisions,
and the accuracy of the mirror's elementary functions."""
import math

import numpy as np
import pytest

import oracle_lib

DEFAULT = dict(mu_a=2.0, mu_s=20.0, shells=101, microns_per_shell=50.0)
# P(u01 <= 0.1f): u01 = (m + 1) 2^-24, m = 0 .. 2^24 - 1 (photon.c:46)
P_SURVIVE = math.floor(float(np.float32(0.1)) * 2**24) / 2**24


def cadence(albedo, cycles=12):
    """Lengths of the weight cycles of photon.c:32,45-49 in double: from 1
    down to the 0.001f cutoff, then from cutoff/0.1f down again, ..."""
    cut, surv = float(np.float32(0.001)), float(np.float32(0.1))
    w, out = 1.0, []
    for _ in range(cycles):
        n = 0
        while w >= cut:
            w *= albedo
            n += 1
        out.append(n)
        w /= surv
    return out


def test_weight_cadence(orc):
    """w_k = a^k until the weight drops below 0.001 (photon.c:32,45): at
    default physics the first roulette follows absorption 73 and the later
    ones follow cycles of 24 absorptions (a 25 every few cycles, as the
    cutoff weight / 0.1 drifts), so every walk length is a cumulative cycle
    sum."""
    p = orc.params(**DEFAULT)
    a = 20.0 / 22.0
    lens = cadence(a)
    assert lens[:4] == [73, 24, 24, 24] and 25 in lens
    allowed = set(np.cumsum(lens).tolist())
    counts = []
    survivor = None
    for photon in range(2000):
        sh, w, n = orc.trace64(p, 7, photon)
        counts.append(n)
        assert w[0] == 1.0
        np.testing.assert_allclose(w[:73], a ** np.arange(73), rtol=1e-12)
        if n > 97 and survivor is None:
            survivor = w
    assert all(n in allowed for n in counts)
    assert survivor is not None
    np.testing.assert_allclose(survivor[73], a**73 / float(np.float32(0.1)), rtol=1e-12)
    assert survivor[72] >= 0.001 and survivor[73] * 0.1 < 0.001


@pytest.mark.parametrize("which", ["walk64", "mirror32"])
@pytest.mark.parametrize("sampler", [0, 1])
def test_mean_interactions_and_energy(orc, which, sampler):
    """Mean absorptions 73 + 24 p / (1 - p); absorbed energy per photon -> 1
    (sum over shells of (1 - a) w, photon.c:30, with an unbiased roulette)."""
    p = orc.params(**DEFAULT)
    N = 1 << 16
    if which == "walk64":
        heat, _, n = orc.walk64(p, 123, N, sampler=sampler)
    else:
        t, n = orc.mirror32(p, 123, N, sampler=sampler)
        heat, _ = oracle_lib.tally_heat(t, 101, 2.0, 20.0)
    mean = n / N
    expect = 73 + 24 * P_SURVIVE / (1 - P_SURVIVE)
    se = 24 * math.sqrt(P_SURVIVE) / (1 - P_SURVIVE) / math.sqrt(N)
    assert abs(mean - expect) < 5 * se
    assert abs(heat.sum() / N - 1.0) < 1e-4


def test_first_step_shell_law(orc):
    """With one absorption per photon the deposit lies at distance t ~ Exp(1)
    mean free paths along +z (photon.c:12-24), so shell s (of width 1/spm mfp)
    has probability exp(-s/spm) - exp(-(s+1)/spm); chi-square over shells."""
    p = orc.params(**DEFAULT)
    N = 1 << 18
    heat, heat2, n = orc.walk64(p, 99, N, max_k=1)
    assert n == N
    oma = 1.0 - 20.0 / 22.0
    counts = heat / oma
    np.testing.assert_allclose(heat2, oma * heat, rtol=1e-12)
    spm = 1e4 / 50.0 / 22.0
    s = np.arange(101)
    prob = np.exp(-s / spm) - np.exp(-(s + 1) / spm)
    prob[-1] = np.exp(-100 / spm)
    expect = N * prob
    keep = expect > 5
    chi2 = float((((counts - expect) ** 2) / expect)[keep].sum())
    dof = int(keep.sum()) - 1
    assert chi2 < dof + 6 * math.sqrt(2 * dof)


@pytest.mark.parametrize("sampler", [0, 1])
def test_direction_moments(orc, sampler):
    """Isotropic scattering (g = 0): unit vectors with E[u] = 0, E[u^2] = 1/3,
    E[uv] = 0 — the law of the reference's rejection method (photon.c:34-43)."""
    d = np.concatenate([orc.directions64(5, ph, sampler, 4096) for ph in range(64)])
    n = d.shape[0]
    np.testing.assert_allclose(np.einsum("ij,ij->i", d, d), 1.0, atol=1e-12)
    se1 = math.sqrt(1 / 3 / n)
    se2 = math.sqrt((1 / 5 - 1 / 9) / n)
    assert np.all(np.abs(d.mean(axis=0)) < 5 * se1)
    assert np.all(np.abs((d * d).mean(axis=0) - 1 / 3) < 5 * se2)
    cross = np.array([(d[:, 0] * d[:, 1]).mean(), (d[:, 0] * d[:, 2]).mean(), (d[:, 1] * d[:, 2]).mean()])
    assert np.all(np.abs(cross) < 5 * math.sqrt(1 / 15 / n))


@pytest.mark.parametrize("sampler", [0, 1])
@pytest.mark.parametrize("phys", [DEFAULT, dict(mu_a=10.0, mu_s=10.0, shells=51, microns_per_shell=100.0),
                                  dict(mu_a=0.5, mu_s=50.0, shells=201, microns_per_shell=20.0),
                                  dict(mu_a=19.0, mu_s=1.0, shells=51, microns_per_shell=100.0)])
def test_mirror32_agrees_with_walk64(orc, sampler, phys):
    """The FP32 mirror and the FP64 oracle walk the same uniforms: the same
    roulette decisions (identical absorption counts) and per-shell heat equal
    to rounding level, except where a deposit lands within rounding of a shell
    boundary."""
    p = orc.params(**phys)
    N = 1 << 15
    h64, _, n64 = orc.walk64(p, 31337, N, sampler=sampler)
    t, n32 = orc.mirror32(p, 31337, N, sampler=sampler)
    h32, _ = oracle_lib.tally_heat(t, phys["shells"], phys["mu_a"], phys["mu_s"])
    assert n32 == n64
    assert abs(h32.sum() / h64.sum() - 1) < 1e-6
    big = h64 > 1e-3 * h64.max()
    assert np.all(np.abs(h32[big] / h64[big] - 1) < 2e-3)


def test_mirror_math_accuracy(orc):
    """Own log and sin/cos(2 pi a) over every uniform the walk draws."""
    log_ulp, sc_abs, sc_ulp = orc.math32_errors()
    assert log_ulp <= 1.0
    assert sc_abs <= 1.2e-7
    assert sc_ulp <= 2.0


def test_derived_constants_match_reference_arithmetic(orc):
    """albedo in float and shells_per_mfp via a double quotient, photon.c:8-9."""
    a, spm = orc.derived32(orc.params(**DEFAULT))
    assert a == np.float32(np.float32(20.0) / np.float32(22.0))
    assert spm == np.float32(1e4 / 50.0 / float(np.float32(22.0)))


@pytest.mark.parametrize("sampler", [0, 1])
def test_mirror_table_mode_equals_own_math(orc, sampler):
    """The mirror's table mode (used for the kernel's hardware math) looks up
    entry r >> 8: fed the own polynomials' values and the correctly rounded
    sin(theta) it reproduces the own-math mirror bit for bit."""
    u64 = (np.arange(1 << 24, dtype=np.float64) + 1.0) * 2.0 ** -24
    u = u64.astype(np.float32)
    lg, sn, cs = orc.eval_math32(u)
    # mu = 2u - 1 and 1 - mu^2 are exact in float64; one rounding to float32
    # is the FMA's, and numpy's float32 sqrt is correctly rounded
    mu = 2.0 * u64 - 1.0
    sp = np.sqrt((1.0 - mu * mu).astype(np.float32))
    tables = (-lg, sn, cs, sp)
    p = orc.params()
    a, na = orc.mirror32(p, 77, 4096, sampler=sampler)
    b, nb = orc.mirror32(p, 77, 4096, sampler=sampler, tables=tables)
    np.testing.assert_array_equal(a, b)
    assert na == nb
    # a table that differs from the polynomials changes the tallies (the
    # rejection sampler does not read sin(theta))
    c, _ = orc.mirror32(p, 77, 4096, sampler=sampler, tables=(-lg * np.float32(1.001), sn, cs, sp))
    assert not np.array_equal(a, c)
    d, _ = orc.mirror32(p, 77, 4096, sampler=sampler, tables=(-lg, sn, cs, sp * np.float32(1.001)))
    assert np.array_equal(a, d) == (sampler == 1)


@pytest.mark.parametrize("phys,N", [(DEFAULT, 1 << 20), (dict(mu_a=10.0, mu_s=10.0, shells=51,
                                                              microns_per_shell=100.0), 1 << 20),
                                    (dict(mu_a=0.5, mu_s=50.0, shells=201, microns_per_shell=20.0), 1 << 16)])
def test_cadence_moments_closed_form(orc, phys, N):
    """The exact per-photon means of absorptions, absorbed energy and summed
    heat2 images from the binary32 weight cadence (oracle_lib.cadence_moments,
    used to check the full-size GPU run) agree with the FP32 mirror within
    5 standard errors.  At default physics the band is narrow enough that a
    cadence off by one absorption, or energy without the (1 - a) factor,
    falls outside it."""
    (mk, vk), (me, ve), (mq, vq) = oracle_lib.cadence_moments(phys["mu_a"], phys["mu_s"])
    t, n = orc.mirror32(orc.params(**phys), 4321, N)
    S = phys["shells"]
    assert abs(n / N - mk) < 5 * math.sqrt(vk / N)
    oma = 1.0 - phys["mu_s"] / (phys["mu_s"] + phys["mu_a"])
    w = sum(int(a) + (int(b) << 64) for a, b in zip(t[0:2 * S:2], t[1:2 * S:2]))
    assert abs(oma * w / 2.0**33 / N - me) < 5 * math.sqrt(ve / N)
    q = sum(int(a) + (int(b) << 64) for a, b in zip(t[2 * S::2], t[2 * S + 1::2]))
    assert abs(q / N - mq) < 5 * math.sqrt(vq / N)
    if phys is DEFAULT:
        assert abs((mk + 1) - n / N) > 5 * math.sqrt(vk / N)
        assert abs(w / 2.0**33 / N - me) > 5 * math.sqrt(ve / N)
