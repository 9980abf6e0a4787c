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


WAIT_S = 60.0


def wait(srv, rid, timeout=WAIT_S):
    """srv.wait with a deadline: a request that does not finish fails its
    test instead of holding the suite."""
    res = srv.wait(rid, timeout)
    assert res is not None, f"request {rid} not finished after {timeout} s"
    return res


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
        got = [wait(srv, i) for i in ids]
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
        got = [wait(srv, i) for i in ids]
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
        t, rs = wait(srv, rid)
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
        a = wait(srv, srv.submit(tmc.Params(photons=50000), 1))
        time.sleep(0.2)  # the kernel leaves after 10 ms without work
        b = wait(srv, srv.submit(tmc.Params(photons=50000), 2))
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
        got = [wait(srv, i) for i in ids]
        a = wait(srv, first)
        st = srv.stats()
    assert st["chunks_published"] > 1 << 15
    np.testing.assert_array_equal(a[0], alone(tmc, *long_req))
    for (p, s), (t, _) in zip(burst, got):
        np.testing.assert_array_equal(t, alone(tmc, p, s))


def test_single_runs_while_serving(tmc):
    """With two workgroup slots per compute unit reserved, a single run on
    the same device completes while the serving kernel walks a long request
    (seconds: one chunk of 64 photons per lane at albedo 0.9999), and
    matches the same run on an idle device."""
    long_req = (tmc.Params(mu_a=0.01, mu_s=100.0, shells=101, photons=1 << 12), 9)
    small = (tmc.Params(photons=20000), 3)
    with tmc.Server(max_shells=101, max_photons_per_lane=64, chunk_work=1 << 40, max_chunk_scale=1,
                    reserve_per_cu=2) as srv:
        rid = srv.submit(*long_req)
        time.sleep(0.05)
        during = alone(tmc, *small)
        overlapped = srv.stats()["completed"] == 0
        a = wait(srv, rid)
    assert overlapped, "the single run waited for the serving kernel"
    np.testing.assert_array_equal(during, alone(tmc, *small))
    np.testing.assert_array_equal(a[0], alone(tmc, *long_req))


def test_two_servers_share_the_device(tmc):
    """Two servers open on one device at once, requests alternating between
    them: every result is exact (their kernels may take turns: a resident
    kernel holds its hardware queue), and a closed server refuses new
    requests."""
    reqs = mixed_requests(tmc, 24, seed=41)
    a, b = tmc.Server(), tmc.Server(max_shells=401)
    try:
        ids = [(srv, srv.submit(p, s)) for i, (p, s) in enumerate(reqs) for srv in [(a, b)[i % 2]]]
        got = [wait(srv, rid) for srv, rid in ids]
    finally:
        a.close(timeout=WAIT_S)
        b.close(timeout=WAIT_S)
    for (p, s), (t, _) in zip(reqs, got):
        np.testing.assert_array_equal(t, alone(tmc, p, s))
    with pytest.raises(tmc.TmcError):
        a.submit(tmc.Params(), 1)


def test_other_work_while_a_server_holds_a_long_request(tmc):
    """While one server's kernel (two workgroup slots per compute unit
    reserved) walks a long request (one chunk of 128 photons per lane at
    albedo 0.9999: a single wave with about 10^7 steps to walk, seconds), a
    second server opens and closes, a batch plan is created, run and
    destroyed, and a single run completes: none of them waits for the
    resident kernel (pinned-memory frees are deferred until no serving
    kernel runs, and nothing synchronises the whole device or a stream the
    resident kernel may share a hardware queue with), and all results are
    exact."""
    long_req = (tmc.Params(mu_a=0.01, mu_s=100.0, shells=101, photons=1 << 13), 77)
    small = (tmc.Params(photons=20000), 3)
    sets = [(tmc.Params(photons=5000 + 100 * i, shells=51), 200 + i) for i in range(6)]
    with tmc.Server(max_shells=101, max_photons_per_lane=128, chunk_work=1 << 40, max_chunk_scale=1,
                    reserve_per_cu=2) as a:
        rid = a.submit(*long_req)
        t0 = time.perf_counter()
        time.sleep(0.05)
        b = tmc.Server(max_shells=101)
        b.close(timeout=WAIT_S)
        t1 = time.perf_counter()
        with tmc.BatchPlan(sets) as plan:
            batch, _ = plan.run()
        single = alone(tmc, *small)
        t2 = time.perf_counter()
        done_meanwhile = a.stats()["completed"]
        got_a = wait(a, rid, 120.0)
        t3 = time.perf_counter()
    assert done_meanwhile == 0, (f"the long request ended before the other work (second server {t1 - t0:.2f} s, "
                                 f"batch + single {t2 - t1:.2f} s, long request {t3 - t0:.2f} s)")
    np.testing.assert_array_equal(single, alone(tmc, *small))
    for (p, s), t in zip(sets, tmc.split_tallies(sets, batch)):
        np.testing.assert_array_equal(t, alone(tmc, p, s))
    np.testing.assert_array_equal(got_a[0], alone(tmc, *long_req))


def test_close_finishes_submitted(tmc):
    srv = tmc.Server()
    for i in range(5):
        srv.submit(tmc.Params(photons=100000), i)
    srv.close(timeout=WAIT_S)


def test_open_loop_load_exact(tmc):
    """A short Poisson stream: every result exact, latencies recorded."""
    reqs = mixed_requests(tmc, 60, seed=17)
    with tmc.Server() as srv:
        recs, wall = tmc.serve_load.run_open_loop(srv, reqs, rate_per_s=400.0)
    lat = np.array([r["stats"].latency_ms for r in recs])
    assert np.all(lat > 0)
    for (p, s), r in zip(reqs, recs):
        np.testing.assert_array_equal(r["tally"], alone(tmc, p, s))


# --- cancellation, deadlines, bounded close, ring reuse by protocol ---------

# A request of 16,384 one-photon-per-lane chunks at albedo 0.9999: more
# chunks than the device has waves (about 8,000), each about 10^5 steps, so
# part of it is still unpublished for tens of milliseconds after it starts.
LONG = (dict(mu_a=0.01, mu_s=100.0, shells=101, photons=1 << 20), 4242)


def long_request(tmc):
    return tmc.Params(**LONG[0]), LONG[1]


def test_cancel_mid_flight_frees_the_slot(tmc):
    """Two request slots: a long request is cancelled while part of it is
    still unpublished; its slot is reused by the requests after it, whose
    results (and the one submitted beside it) are exact."""
    small = [(tmc.Params(photons=30000 + 1000 * i, shells=51 + 50 * (i % 3)), 300 + i) for i in range(5)]
    with tmc.Server(max_requests=2, max_shells=201, max_photons_per_lane=1) as srv:
        rid = srv.submit(*long_request(tmc))
        first = srv.submit(*small[0])
        time.sleep(0.03)
        assert srv.poll(rid) is None
        srv.cancel(rid)
        later = [srv.submit(p, s) for p, s in small[1:]]  # need the cancelled request's slot
        got = [wait(srv, i) for i in [first] + later]
        with pytest.raises(tmc.TmcError) as e:
            wait(srv, rid)
        assert e.value.status == tmc.E_CANCELLED
        st = srv.stats()
    assert st["cancelled"] == 1 and st["completed"] == len(small)
    for (p, s), (t, _) in zip(small, got):
        np.testing.assert_array_equal(t, alone(tmc, p, s))


def test_wait_gives_up_then_cancel(tmc):
    """A host-side wait that gives up (timeout) leaves the request running;
    cancelling it then ends it within the drain of its claimed chunks."""
    with tmc.Server(max_shells=101, max_photons_per_lane=1) as srv:
        rid = srv.submit(*long_request(tmc))
        assert srv.wait(rid, timeout=0.02) is None
        srv.cancel(rid)
        t0 = time.perf_counter()
        with pytest.raises(tmc.TmcError) as e:
            wait(srv, rid)
        assert e.value.status == tmc.E_CANCELLED
        assert time.perf_counter() - t0 < 30.0
        after = wait(srv, srv.submit(tmc.Params(photons=40000), 5))
    np.testing.assert_array_equal(after[0], alone(tmc, tmc.Params(photons=40000), 5))


def test_deadline_cancels_late_request(tmc):
    p, s = long_request(tmc)
    with tmc.Server(max_shells=101, max_photons_per_lane=1) as srv:
        late = srv.submit(p, s, deadline=0.05)
        ok = srv.submit(tmc.Params(photons=20000), 8, deadline=60.0)
        with pytest.raises(tmc.TmcError) as e:
            wait(srv, late)
        assert e.value.status == tmc.E_DEADLINE
        got = wait(srv, ok)
        st = srv.stats()
    assert st["expired"] == 1
    np.testing.assert_array_equal(got[0], alone(tmc, tmc.Params(photons=20000), 8))


def test_bounded_close_cancels_and_returns(tmc):
    """close with a time limit: the unfinished requests are cancelled when it
    runs out, their claimed chunks drain and the kernel leaves."""
    srv = tmc.Server(max_shells=101, max_photons_per_lane=1)
    srv.submit(*long_request(tmc))
    for i in range(3):
        srv.submit(tmc.Params(photons=50000), 60 + i)
    time.sleep(0.05)
    t0 = time.perf_counter()
    srv.close(timeout=5.0)
    elapsed = time.perf_counter() - t0
    print(f"bounded close: {elapsed:.3f} s", flush=True)
    assert elapsed < 12.0
    with pytest.raises(tmc.TmcError):
        srv.submit(tmc.Params(), 1)


def test_tiny_ring_wraps_many_times(tmc):
    """A 64-entry ring carries about 20,000 chunks (300+ laps): each slot is
    rewritten only after its previous entry was read, and every result is
    exact."""
    reqs = [(tmc.Params(photons=65536, shells=51 + 50 * (i % 4)), 500 + i) for i in range(20)]
    with tmc.Server(ring_chunks=64, max_photons_per_lane=1, max_chunk_scale=1) as srv:
        ids = [srv.submit(p, s) for p, s in reqs]
        got = [wait(srv, i) for i in ids]
        st = srv.stats()
    assert st["chunks_published"] >= 20 * 1024 > 300 * 64
    for (p, s), (t, _) in zip(reqs, got):
        np.testing.assert_array_equal(t, alone(tmc, p, s))


def test_served_results_within_tolerance_of_fp64_oracle(tmc, orc):
    """Per-request accuracy at the single run's tolerance bar: each served result's
    heat, shell by shell, within max(3 sigma, 1e-5 relative) of the FP64
    oracle on the same stream, for requests of four physics served together
    (sigma: the larger of the heat2 estimate and the standard error over 8
    photon batches of the oracle)."""
    cases = [(dict(), 1 << 18, 11),
             (dict(mu_a=10.0, mu_s=10.0, shells=51, microns_per_shell=100.0), 1 << 18, 12),
             (dict(mu_a=0.5, mu_s=50.0, shells=201, microns_per_shell=20.0), 1 << 15, 13),
             (dict(mu_a=1.0, mu_s=60.0, shells=401, microns_per_shell=10.0), 1 << 16, 14)]
    with tmc.Server() as srv:
        ids = [srv.submit(tmc.Params(photons=N, **phys), seed) for phys, N, seed in cases]
        got = [wait(srv, i) for i in ids]
    for (phys, N, seed), (tally, _) in zip(cases, got):
        p = tmc.Params(photons=N, **phys)
        heat, _ = tmc.tally_to_heat(p, tally)
        B = 8
        parts = [orc.walk64(orc.params(**phys), seed, N // B, first=b * (N // B)) for b in range(B)]
        batches = np.array([h for h, _, _ in parts])
        oh = batches.sum(axis=0)
        oh2 = np.sum([h2 for _, h2, _ in parts], axis=0)
        sigma = np.maximum(np.sqrt(np.maximum(oh2 - oh * oh / N, 0.0)), np.sqrt(B) * batches.std(axis=0, ddof=1))
        tol = np.maximum(3.0 * sigma, 1e-5 * np.abs(oh))
        bad = np.nonzero(np.abs(heat - oh) > tol)[0]
        assert bad.size == 0, (phys, bad, heat[bad], oh[bad], tol[bad])
