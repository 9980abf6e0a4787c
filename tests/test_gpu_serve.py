"""Serving (continuous batching on a resident persistent kernel): each
result equals the same request run alone, bit for bit, whether requests
arrive together, staggered while others run, or wait for memory."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def mixed_requests(tmc, n, seed):
    reqs = tmc.serve_load.draw_requests(n, seed=seed)
    # keep the suite quick: shrink the heavy tail
    return [(tmc.Params(p.mu_a, p.mu_s, p.shells, p.microns_per_shell, max(64, p.photons // 64)), s)
            for p, s in reqs]


def alone(tmc, p, s, sampler=0):
    return tmc.run_raw(p, s, sampler=sampler)[0]


@pytest.mark.parametrize("sampler", [0, 1])
def test_alone_equals_in_the_mix(tmc, sampler):
    reqs = mixed_requests(tmc, 40, seed=3 + sampler)
    with tmc.Server(sampler=sampler) as srv:
        ids = []
        for i, (p, s) in enumerate(reqs):
            ids.append(srv.submit(p, s))
            if i % 10 == 9:
                time.sleep(0.02)  # later requests arrive while earlier ones run
        got = [srv.wait(i) for i in ids]
        st = srv.stats()
    for (p, s), (t, rs) in zip(reqs, got):
        np.testing.assert_array_equal(t, alone(tmc, p, s, sampler))
        assert rs.photons == p.photons and rs.latency_ms >= 0 and rs.done_s >= rs.admit_s >= rs.submit_s
    assert st["completed"] == len(reqs)


def test_admission_waits_for_memory(tmc):
    """A tally arena with room for two 101-shell requests at a time: every
    request still completes, in arrival order of admission, with exact
    results; a request whose tally can never fit is refused."""
    reqs = [(tmc.Params(photons=20000 + 1000 * i), 50 + i) for i in range(8)]
    with tmc.Server(max_requests=2, tally_words=2 * 4 * 101, max_shells=101) as srv:
        ids = [srv.submit(p, s) for p, s in reqs]
        got = [srv.wait(i) for i in ids]
        with pytest.raises(tmc.TmcError) as e:
            srv.submit(tmc.Params(shells=401), 1)
        assert e.value.status == -1  # more shells than max_shells
    admits = [rs.admit_s for _, rs in got]
    assert admits == sorted(admits)
    for (p, s), (t, _) in zip(reqs, got):
        np.testing.assert_array_equal(t, alone(tmc, p, s))
    with tmc.Server(max_requests=2, tally_words=100, max_shells=101) as srv:
        with pytest.raises(tmc.TmcError) as e:
            srv.submit(tmc.Params(), 1)
        assert e.value.status == -4


def test_zero_photons_bad_params_and_poll(tmc):
    with tmc.Server() as srv:
        rid = srv.submit(tmc.Params(photons=0), 1)
        t, rs = srv.wait(rid)
        assert not t.any() and rs.photons == 0
        with pytest.raises(tmc.TmcError):
            srv.submit(tmc.Params(mu_a=0.0), 1)
        with pytest.raises(tmc.TmcError) as e:
            srv.submit(tmc.Params(photons=10), 1, first_photon=(1 << 63) - 5)
        assert e.value.status == -5
        rid = srv.submit(tmc.Params(photons=200000), 7)
        res = None
        for _ in range(20000):
            res = srv.poll(rid)
            if res is not None:
                break
            time.sleep(1e-3)
        assert res is not None
        np.testing.assert_array_equal(res[0], alone(tmc, tmc.Params(photons=200000), 7))


def test_kernel_relaunch_after_idle(tmc):
    with tmc.Server(idle_ms=10) as srv:
        a = srv.wait(srv.submit(tmc.Params(photons=50000), 1))
        time.sleep(0.2)  # the kernel leaves after 10 ms without work
        b = srv.wait(srv.submit(tmc.Params(photons=50000), 2))
        st = srv.stats()
    assert st["kernel_launches"] >= 2
    np.testing.assert_array_equal(a[0], alone(tmc, tmc.Params(photons=50000), 1))
    np.testing.assert_array_equal(b[0], alone(tmc, tmc.Params(photons=50000), 2))


def test_burst_past_ring_while_waves_dry(tmc):
    """Most waves sit dry for longer than idle_ms while one long request
    runs; then a burst publishes more chunks than the ring holds.  Every
    ring index a wave took while dry must still be claimed (a wave that left
    on its own with its ticket stranded that chunk once the ring wrapped)."""
    long_req = (tmc.Params(mu_a=0.01, mu_s=100.0, shells=101, photons=64), 5)
    burst = [(tmc.Params(photons=65536, shells=101), 100 + i) for i in range(40)]
    with tmc.Server(idle_ms=10, ring_chunks=1 << 15, max_photons_per_lane=1, max_shells=101) as srv:
        first = srv.submit(*long_req)
        time.sleep(0.05)
        ids = [srv.submit(p, s) for p, s in burst]
        got = [srv.wait(i) for i in ids]
        a = srv.wait(first)
        st = srv.stats()
    assert st["chunks_published"] > 1 << 15
    np.testing.assert_array_equal(a[0], alone(tmc, *long_req))
    for (p, s), (t, _) in zip(burst, got):
        np.testing.assert_array_equal(t, alone(tmc, p, s))


def test_single_runs_while_serving(tmc):
    """A single run on the same device while the serving kernel is resident
    with a long request completes (in the compute units the server leaves
    free) and matches the same run on an idle device."""
    long_req = (tmc.Params(mu_a=0.01, mu_s=100.0, shells=101, photons=64), 9)
    small = (tmc.Params(photons=20000), 3)
    with tmc.Server(max_shells=101) as srv:
        rid = srv.submit(*long_req)
        during = alone(tmc, *small)
        a = srv.wait(rid)
    np.testing.assert_array_equal(during, alone(tmc, *small))
    np.testing.assert_array_equal(a[0], alone(tmc, *long_req))


def test_two_servers_share_the_device(tmc):
    """Two servers, each with its resident kernel, on one device at once:
    neither waits on the other (no copies while a kernel runs), every result
    is exact, and a closed server refuses new requests."""
    reqs = mixed_requests(tmc, 24, seed=41)
    a, b = tmc.Server(), tmc.Server(max_shells=401)
    try:
        ids = [(srv, srv.submit(p, s)) for i, (p, s) in enumerate(reqs) for srv in [(a, b)[i % 2]]]
        got = [srv.wait(rid) for srv, rid in ids]
    finally:
        a.close()
        b.close()
    for (p, s), (t, _) in zip(reqs, got):
        np.testing.assert_array_equal(t, alone(tmc, p, s))
    with pytest.raises(tmc.TmcError):
        a.submit(tmc.Params(), 1)


def test_close_finishes_submitted(tmc):
    srv = tmc.Server()
    for i in range(5):
        srv.submit(tmc.Params(photons=100000), i)
    srv.close()


def test_open_loop_load_exact(tmc):
    """A short Poisson stream: every result exact, latencies recorded."""
    reqs = mixed_requests(tmc, 60, seed=17)
    with tmc.Server() as srv:
        recs, wall = tmc.serve_load.run_open_loop(srv, reqs, rate_per_s=400.0)
    lat = np.array([r["stats"].latency_ms for r in recs])
    assert np.all(lat > 0)
    for (p, s), r in zip(reqs, recs):
        np.testing.assert_array_equal(r["tally"], alone(tmc, p, s))
