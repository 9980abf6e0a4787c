"""Open-loop load for the serving mode: Poisson arrivals of simulation
requests with the sweep's physics (mu_a log-uniform [0.05, 10]/cm, mu_s
log-uniform [10, 200]/cm, shells in {51, 101, 201, 401}, width uniform
[10, 100] um) and heavy-tailed photon counts (Pareto alpha = 1.5, x_m = 2^14,
clipped to [2^14, 2^24]).  The offered load is a stated fraction of a
measured capacity in absorptions per second."""
import threading
import time

import numpy as np

from ._lib import Params, expected_interactions
from .sweep import SHELL_CHOICES

X_M = 2**14
PHOTONS_MAX = 2**24
ALPHA = 1.5


def draw_requests(n, seed=0):
    """n requests: [(Params, seed)].  Request i is a function of (seed, i)
    alone: five uniforms per request, drawn row by row, so the first m of n
    requests are the m requests (an offered rate that differs slightly
    between runs keeps the same list up to its length)."""
    u = np.random.default_rng(seed).uniform(0.0, 1.0, (n, 5))
    mu_a = np.exp(np.log(0.05) + u[:, 0] * (np.log(10.0) - np.log(0.05)))
    mu_s = np.exp(np.log(10.0) + u[:, 1] * (np.log(200.0) - np.log(10.0)))
    shells = np.asarray(SHELL_CHOICES)[np.minimum((u[:, 2] * len(SHELL_CHOICES)).astype(np.int64),
                                                  len(SHELL_CHOICES) - 1)]
    width = 10.0 + u[:, 3] * 90.0
    # Pareto(ALPHA, X_M) by inversion of 1 - u in (0, 1]
    photons = np.clip(X_M / (1.0 - u[:, 4]) ** (1.0 / ALPHA), X_M, PHOTONS_MAX).astype(np.int64)
    return [(Params(float(np.float32(mu_a[i])), float(np.float32(mu_s[i])), int(shells[i]),
                    float(np.float32(width[i])), int(photons[i])), 1_000_000 + i) for i in range(n)]


def mean_request_work(requests):
    return float(np.mean([p.photons * expected_interactions(p) for p, _ in requests]))


def run_open_loop(server, requests, rate_per_s, seed=1, progress=None, timeout_s=300.0):
    """Submits `requests` at Poisson arrival times with `rate_per_s` and
    waits for every result.  Returns per-request records (dict) and the wall
    time from the first arrival to the last completion."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, len(requests))
    arrivals = np.cumsum(gaps) - gaps[0]
    ids = [None] * len(requests)
    submitted = threading.Event()
    records = [None] * len(requests)
    t0 = time.perf_counter()

    def submitter():
        for i, (p, s) in enumerate(requests):
            delay = t0 + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids[i] = server.submit(p, s)
        submitted.set()

    th = threading.Thread(target=submitter)
    th.start()
    for i in range(len(requests)):
        while ids[i] is None:
            time.sleep(1e-4)
        res = server.wait(ids[i], timeout=timeout_s)
        if res is None:
            raise TimeoutError(f"request {i} not finished after {timeout_s} s: {server.stats()}")
        tally, st = res
        records[i] = {"tally": tally, "stats": st}
        if progress and (i + 1) % 2000 == 0:
            progress(f"serve: {i + 1}/{len(requests)} requests done, {time.perf_counter() - t0:.1f} s")
    th.join()
    wall = time.perf_counter() - t0
    return records, wall
