"""tiny_mc_gpu_run_noalloc: the single run with its device words in a static
scratch array instead of an allocation.  Called through the C ABI, it must
give the tallies and counts of tiny_mc_gpu_run_ex bit for bit, on calls in a
row whose sizes shrink (the array keeps the previous call's words, which
each call clears first), and refuse more shells than the array holds."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def call(tmc, entry, params, seed, sampler=0, precision=0):
    from tiny_mc_amd._lib import _Options, _Stats

    lib = tmc.load()
    tally = np.full(4 * params.shells, 0xDEADBEEF, dtype=np.uint64)
    opt = _Options(sampler, 0, 0, 0, 0, 0, precision, 0)
    st = _Stats()
    c = params._c()
    rc = getattr(lib, entry)(ctypes.byref(c), seed, ctypes.byref(opt),
                             tally.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(st))
    return rc, tally, st


def check_same(tmc, params, seed, **kw):
    print(f"noalloc vs run_ex: shells={params.shells} photons={params.photons} seed={seed} {kw}", flush=True)
    rc0, t0, s0 = call(tmc, "tiny_mc_gpu_run_noalloc", params, seed, **kw)
    rc1, t1, s1 = call(tmc, "tiny_mc_gpu_run_ex", params, seed, **kw)
    assert rc0 == 0 and rc1 == 0
    np.testing.assert_array_equal(t0, t1)
    assert (s0.photons, s0.interactions, s0.truncated) == (s1.photons, s1.interactions, s1.truncated)
    assert t0.any()


def test_noalloc_equals_run_ex_on_shrinking_calls(tmc):
    # largest size the scratch holds, then smaller ones: stale words of the
    # earlier call must not leak into the later one
    check_same(tmc, tmc.Params(mu_a=0.5, mu_s=50.0, shells=tmc.NOALLOC_MAX_SHELLS, microns_per_shell=1.0,
                               photons=1 << 16), 3)
    check_same(tmc, tmc.Params(photons=1 << 20), 12345)
    check_same(tmc, tmc.Params(shells=51, photons=5000), 7, sampler=1)
    check_same(tmc, tmc.Params(shells=2, photons=64), 9)


def test_noalloc_precision_variant(tmc):
    check_same(tmc, tmc.Params(photons=1 << 14), 21, precision=3)


def test_noalloc_refuses_more_shells_than_its_scratch(tmc):
    p = tmc.Params(shells=tmc.NOALLOC_MAX_SHELLS + 1, photons=64)
    rc, t, _ = call(tmc, "tiny_mc_gpu_run_noalloc", p, 1)
    assert rc == -7
    assert (t == 0xDEADBEEF).all()  # nothing written
    lib = tmc.load()
    assert b"no-allocation" in lib.tiny_mc_gpu_strerror(-7)
    # the same size still runs through the allocating entry point
    rc, t, _ = call(tmc, "tiny_mc_gpu_run_ex", p, 1)
    assert rc == 0 and t.any()
