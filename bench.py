"""Benchmark of the MI355X photon-transport path (one GPU).

A step is one launch of the walk over 2^30 photons at default physics
(mu_a = 2/cm, mu_s = 20/cm, 101 shells of 50 um, weight cutoff 0.001,
roulette 0.1), through the C ABI call tiny_mc_gpu_run_ex: clear the tallies,
run the persistent kernel, copy the 128-bit tallies back.  There is no input
data: photons are generated from (seed, photon index) on the device.

Prints one JSON line: whole-job photons/s, the roofline of the walk kernel
(algorithmic FP32 FLOPs of the reference algorithm, 40.5 per absorption
event counted from photon.c:21-48, over the kernel's average duration
measured with HIP events on the library's stream), the same walk with the
reference's own rejection sampler (photon.c:34-43), and the reference CPU
loop timed on this machine's host cores.

More legs of the metric follow, each once:
  sweep   4096 parameter sets (mu_a log-uniform [0.05,10]/cm, mu_s
          log-uniform [10,200]/cm, shells {51,101,201,401}, width U[10,100] um,
          photons log-uniform [2^16,2^22], seed = set index) in persistent
          launches with autotuned options (tuned outside the timed runs; three
          timed runs, and two with the default options), beside the same sets
          run one launch per set;
  serve   Poisson arrivals of requests (sweep physics, Pareto(1.5) photons in
          [2^14,2^24]) at half the sweep's measured capacity (latency) and at
          its full capacity (throughput at saturation), with p50/p99 latency,
          aggregate photons/s and a sample of results checked bit for bit
          against the same request run alone; and requests of the saturation
          stream run one launch per request (the plain path, as the
          reference's loop of tiny_mc.c:47-49 per configuration).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--photons N]
                    [--sampler analytic|rejection] [--no-cpu-baseline]
                    [--no-sweep] [--no-serve] [--serve-load F] [--serve-seconds T]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "photons/s at default physics; sweep photons/s + p50/p99 latency at fixed load"
# FP32 FLOPs per absorption event of photon.c:21-48 (add, mul, div, sqrt and
# log 1 each): 2 step, 6 move, 7 radius and shell, 6 tally and weight,
# 19.46 direction by rejection (9 per attempt x 4/pi attempts, 2 for u,
# 6 for v and w), 0.03 roulette.  The algorithmic work the roofline counts.
FLOP_PER_INTERACTION = 40.5
# The analytic sampler does the same work but the direction: 2u - 1 (2),
# 1 - mu^2 (2), sqrt (1), the angle (1), sin and cos (2), v and w (2).
FLOP_PER_INTERACTION_ANALYTIC = 40.5 - 19.46 + 10.0
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 vector peak
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "walk_f32_default_2e30_traffic.txt")
REF_BENCH = os.path.join(ROOT, "oracle", "_ref", "tiny_mc_bench")
REF_SHIPPED = os.path.join(ROOT, "oracle", "_ref", "tiny_mc_shipped")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def ref_rate(exe, seed, photons):
    """K photons/s line of one run of a reference build (tiny_mc.c:53-56)."""
    env = dict(os.environ, TMC_SEED=str(seed), TMC_PHOTONS=str(photons))
    out = subprocess.run([exe], env=env, stdout=subprocess.PIPE, text=True, check=True).stdout
    m = re.search(r"^# ([0-9.]+) K photons per second", out, re.M)
    return float(m.group(1)) * 1e3 if m else None


def progress(msg):
    """One line per finished leg on stderr (long runs stay visibly alive)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(photons_per_proc=1 << 22, max_procs=16):
    """The reference loop (photon() + glibc rand(), unmodified sources built
    by oracle/Makefile) on the host cores: one process per usable core, at
    most max_procs (the host cores this machine allots to one GPU), distinct
    seeds (threads would serialise on rand()'s lock), each about 13 s of
    work.  Sum of photons / wall time."""
    usable = len(os.sched_getaffinity(0))
    procs = max(1, min(max_procs, usable))
    if os.path.exists(REF_BENCH):
        cmd, kind = [REF_BENCH], "reference"
    else:
        # the oracle's byte-identical restatement of the same loop
        cmd, kind = [sys.executable, os.path.abspath(__file__), "--cpu-port-worker"], "port"
    env_key = "TMC_PHOTONS"
    t0 = time.perf_counter()
    running = []
    for i in range(procs):
        env = dict(os.environ, TMC_SEED=str(i + 1), **{env_key: str(photons_per_proc)})
        running.append(subprocess.Popen(cmd, env=env, stdout=subprocess.PIPE, text=True))
    rates, secs = [], []
    for p in running:
        out, _ = p.communicate()
        if p.returncode != 0:
            return None
        m = re.search(r"^# ([0-9.]+) K photons per second", out, re.M)
        if m:
            rates.append(float(m.group(1)) * 1e3)
        m = re.search(r"^# ([0-9.]+) seconds", out, re.M)
        if m:
            secs.append(float(m.group(1)))
    wall = time.perf_counter() - t0
    shipped = ref_rate(REF_SHIPPED, 1, 1 << 19) if os.path.exists(REF_SHIPPED) else None
    return {
        "value": procs * photons_per_proc / wall,
        "unit": "photons/s",
        "cores": procs,
        "kind": kind,
        "cpu_model": cpu_model(),
        "cpu_count": os.cpu_count(),
        "usable_cores": usable,
        "cores_cap": max_procs,
        "cores_cap_reason": "host cores allotted to one GPU on this machine",
        "seconds_per_process": {"min": min(secs), "max": max(secs)} if secs else None,
        "single_core_value": sum(rates) / len(rates) if rates else None,
        "single_core_as_shipped_value": shipped,
        "sample": f"{procs} processes x {photons_per_proc} photons, default physics, seeds 1..{procs}; "
                  + ("unmodified reference tiny_mc.c/photon.c/wtime.c built -std=c11 -O3 -march=x86-64-v3"
                     if kind == "reference" else "oracle restatement of the reference loop (oracle/liboracle.so)"),
    }


def cpu_port_worker():
    """One process of the fallback CPU baseline: the oracle's restatement of
    the reference loop (glibc rand(), FP32), timed like tiny_mc.c:45-56."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    orc = oracle_lib.load()
    photons = int(os.environ["TMC_PHOTONS"])
    t0 = time.perf_counter()
    orc.restate(orc.params(), int(os.environ["TMC_SEED"]), photons)
    dt = time.perf_counter() - t0
    print(f"# {dt:f} seconds\n# {1e-3 * photons / dt:f} K photons per second")


def copy_calibration(nbytes=1 << 30):
    """A plain device copy of known size, for scaling the FETCH_SIZE /
    WRITE_SIZE counters of the walk kernel (run under rocprofv3 --pmc)."""
    import torch

    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 3}))


def rocpd_summary(db_path):
    """Text summary of a rocprofv3 output database (rocpd SQLite): per kernel
    the dispatch count and duration statistics, and per kernel and counter
    the per-dispatch value (summed over hardware instances) averaged over
    dispatches."""
    import sqlite3

    con = sqlite3.connect(db_path)
    lines = [f"# rocpd summary of {os.path.basename(db_path)}"]
    rows = con.execute("select name, count(*), avg(duration), min(duration), max(duration) from kernels "
                       "group by name order by sum(duration) desc").fetchall()
    if rows:
        lines.append("kernel | calls | avg_ns | min_ns | max_ns")
        lines += [f"{n} | {c} | {a:.0f} | {lo} | {hi}" for n, c, a, lo, hi in rows]
    rows = con.execute("select kernel_name, counter_name, dispatch_id, sum(value) from counters_collection "
                       "group by kernel_name, counter_name, dispatch_id").fetchall()
    if rows:
        acc = {}
        for kname, cname, _, v in rows:
            acc.setdefault((kname, cname), []).append(v)
        lines.append("kernel | counter | dispatches | avg_per_dispatch")
        lines += [f"{k} | {c} | {len(v)} | {sum(v) / len(v):.6g}" for (k, c), v in sorted(acc.items())]
    return "\n".join(lines) + "\n"


def sweep_leg(tuned_runs=3, default_runs=2):
    """The batched sweep in persistent launches vs one launch per set (same
    sets).  Options autotuned on a 1/16 sample before the timed runs."""
    import statistics

    import tiny_mc_amd

    sets = tiny_mc_amd.sweep_sets(n=4096, seed=0)
    photons = sum(p.photons for p, _ in sets)
    # warm the multi-set kernel's code object on a small batch
    tiny_mc_amd.batch_run(tiny_mc_amd.sweep_sets(n=8, seed=1, photon_scale=1e-3))
    t0 = time.perf_counter()
    tuned = tiny_mc_amd.batch_autotune(sets, tiny_mc_amd.BatchOptions(), sample_divisor=16)
    tune_s = time.perf_counter() - t0
    tuned.use_graph = True
    progress(f"sweep: autotuned in {tune_s:.1f} s: {tuned}")

    def timed(options, runs):
        out, st = [], None
        with tiny_mc_amd.BatchPlan(sets, options) as plan:
            for _ in range(runs):
                t = time.perf_counter()
                _, st = plan.run()
                out.append(time.perf_counter() - t)
        return out, st

    tuned_s, st = timed(tuned, tuned_runs)
    progress(f"sweep batched (tuned): {', '.join(f'{x:.2f}' for x in tuned_s)} s")
    default_s, _ = timed(tiny_mc_amd.BatchOptions(use_graph=True), default_runs)
    progress(f"sweep batched (defaults): {', '.join(f'{x:.2f}' for x in default_s)} s; now one launch per set")
    t0 = time.perf_counter()
    for p, seed in sets:
        tiny_mc_amd.run_raw(p, seed)
    per_set_s = time.perf_counter() - t0
    med = statistics.median(tuned_s)
    return {
        "sets": len(sets),
        "photons": photons,
        "interactions": st.interactions,
        "seconds": med,
        "seconds_runs": tuned_s,
        "seconds_min": min(tuned_s),
        "photons_per_s": photons / med,
        "interactions_per_s": st.interactions / med,
        "default_options_seconds_runs": default_s,
        "tuned_vs_default": statistics.median(default_s) / med,
        "one_launch_per_set_seconds": per_set_s,
        "one_launch_per_set_photons_per_s": photons / per_set_s,
        "speedup_vs_one_launch_per_set": per_set_s / med,
        "speedup_vs_one_launch_per_set_best": per_set_s / min(tuned_s),
        "chunks": st.chunks,
        "lane_utilisation": st.lane_utilisation,
        "autotune_seconds": tune_s,
        "launch": {"passes": st.passes, "grid": st.grid, "block": st.block, "waves_per_simd": st.waves_per_simd,
                   "copies": st.copies, "chunk_work": st.chunk_work, "hip_graph": True,
                   "options": {"block": tuned.block, "waves_per_simd": tuned.waves_per_simd, "copies": tuned.copies,
                               "chunk_work": tuned.chunk_work}},
    }


def serve_leg(capacity_interactions_per_s, load=0.5, seconds=15.0, check=16, seed=12, max_chunk_scale=0):
    """Open-loop Poisson load at `load` x capacity for about `seconds`."""
    import numpy as np

    import tiny_mc_amd

    probe = tiny_mc_amd.serve_load.draw_requests(4000, seed=11)
    mean_work = tiny_mc_amd.serve_load.mean_request_work(probe)
    rate = load * capacity_interactions_per_s / mean_work
    n = max(50, int(rate * seconds))
    reqs = tiny_mc_amd.serve_load.draw_requests(n, seed=seed)
    with tiny_mc_amd.Server(max_chunk_scale=max_chunk_scale) as srv:
        progress(f"serve: {n} requests at {rate:.4g}/s ({load:.2f} of capacity)")
        recs, wall = tiny_mc_amd.serve_load.run_open_loop(srv, reqs, rate, progress=progress)
        stats = srv.stats()
    lat = np.array([r["stats"].latency_ms for r in recs])
    photons = sum(p.photons for p, _ in reqs)
    inter = sum(r["stats"].interactions for r in recs)
    rng = np.random.default_rng(5)
    sample = rng.choice(len(reqs), size=min(check, len(reqs)), replace=False)
    exact = sum(int(np.array_equal(recs[i]["tally"], tiny_mc_amd.run_raw(*reqs[i])[0])) for i in sample)
    return {
        "offered_load": load,
        "offered_requests_per_s": rate,
        "requests": n,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()),
        "photons_per_s": photons / wall,
        "interactions_per_s": inter / wall,
        "seconds": wall,
        "kernel_launches": stats["kernel_launches"],
        "exact_vs_alone": f"{exact}/{len(sample)}",
    }, reqs, recs


def one_launch_per_request_leg(reqs, count=1500):
    """The plain path on the served requests: each request (a random subset
    of the stream) as one synchronous launch, back to back."""
    import numpy as np

    import tiny_mc_amd

    rng = np.random.default_rng(7)
    idx = rng.choice(len(reqs), size=min(count, len(reqs)), replace=False)
    inter, photons = 0, 0
    t0 = time.perf_counter()
    for i in idx:
        p, seed = reqs[i]
        _, st = tiny_mc_amd.run_raw(p, seed)
        inter += st.interactions
        photons += p.photons
    wall = time.perf_counter() - t0
    return {"requests": len(idx), "seconds": wall, "photons_per_s": photons / wall,
            "interactions_per_s": inter / wall}


def same_work_batch_leg(reqs):
    """The served stream's requests as one batched plan (default options),
    timed like the sweep's plan runs: the batch kernel's rate on the very
    work the server walked, for the served/batch capacity ratio."""
    import tiny_mc_amd

    with tiny_mc_amd.BatchPlan(reqs, tiny_mc_amd.BatchOptions(use_graph=True)) as plan:
        t = time.perf_counter()
        _, st = plan.run()
        wall = time.perf_counter() - t
    photons = sum(p.photons for p, _ in reqs)
    return {"requests": len(reqs), "seconds": wall, "photons_per_s": photons / wall,
            "interactions_per_s": st.interactions / wall, "lane_utilisation": st.lane_utilisation}


def reference_case_leg(sampler, calls=20):
    """The reference's own configuration (params.h defaults: 32,768 photons,
    101 shells): too small to fill the GPU, so the figure is the latency of
    one synchronous call (median of `calls`)."""
    import tiny_mc_amd

    p = tiny_mc_amd.Params(photons=32768)
    total, kern = [], []
    for i in range(calls + 2):
        _, st = tiny_mc_amd.run_raw(p, 12345 + i, sampler=sampler)
        if i >= 2:
            total.append(st.total_ms)
            kern.append(st.kernel_ms)
    return {
        "photons": p.photons,
        "call_ms_p50": float(sorted(total)[len(total) // 2]),
        "kernel_ms_p50": float(sorted(kern)[len(kern) // 2]),
        "photons_per_s_per_call": p.photons / (sorted(total)[len(total) // 2] * 1e-3),
    }


def rejection_leg(params, steps):
    """The same walk with the reference's rejection sampler (photon.c:34-43)."""
    import tiny_mc_amd

    tiny_mc_amd.run_raw(params, 999, sampler=tiny_mc_amd.SAMPLER_REJECTION)
    kernel_ms, inter = [], 0
    t0 = time.perf_counter()
    for i in range(steps):
        _, st = tiny_mc_amd.run_raw(params, 500 + i, sampler=tiny_mc_amd.SAMPLER_REJECTION)
        kernel_ms.append(st.kernel_ms)
        inter += st.interactions
    wall = time.perf_counter() - t0
    k_s = sum(kernel_ms) / len(kernel_ms) * 1e-3
    achieved = FLOP_PER_INTERACTION * inter / steps / k_s / 1e12
    return {"photons_per_s": steps * params.photons / wall, "kernel_ms_avg": k_s * 1e3, "steps": steps,
            "interactions_per_s": inter / wall, "roofline_achieved": achieved, "frac": achieved / PEAK_FP32_TFLOPS}


def read_traffic():
    """HBM bytes per walk launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            m = re.search(r"^traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", f.read(), re.M)
        return float(m.group(1)) if m else None
    except OSError:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--photons", type=int, default=1 << 30)
    ap.add_argument("--sampler", choices=["analytic", "rejection"], default="analytic")
    ap.add_argument("--weight-table", type=int, default=0,
                    help="tmc_options.weight_table: 0 automatic, -1 off, N at most N entries")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-port-worker", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 1 GiB device copy (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--rocpd-summary", metavar="DB", help="print a text summary of a rocprofv3 results database")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-reference-case", action="store_true",
                    help="skip the 32,768-photon call latency (keeps profiled runs to the 2^30-photon launches)")
    ap.add_argument("--serve-load", type=float, default=0.5)
    ap.add_argument("--serve-seconds", type=float, default=12.0)
    ap.add_argument("--saturation-seconds", type=float, default=8.0)
    ap.add_argument("--no-saturation", action="store_true",
                    help="skip the served run at saturation and its one-launch-per-request comparison")
    ap.add_argument("--no-rejection", action="store_true", help="skip the rejection-sampler figure")
    ap.add_argument("--serve-max-chunk-scale", type=int, default=0,
                    help="tmc_serve_options.max_chunk_scale of the served legs (0: the default)")
    ap.add_argument("--serve-capacity", type=float, default=0.0,
                    help="absorptions/s the serving loads are fractions of (0: the sweep's measured rate)")
    args = ap.parse_args()
    if args.rocpd_summary:
        return print(rocpd_summary(args.rocpd_summary), end="")
    if args.cpu_port_worker:
        return cpu_port_worker()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        sys.exit("bench.py: this path runs on one GPU only; --gpus must be 1 "
                 "(photon ranges shard across GPUs with one 2*SHELLS-word sum, not measured here)")

    import tiny_mc_amd

    sampler = tiny_mc_amd.SAMPLER_ANALYTIC if args.sampler == "analytic" else tiny_mc_amd.SAMPLER_REJECTION
    params = tiny_mc_amd.Params(mu_a=2.0, mu_s=20.0, shells=101, microns_per_shell=50.0, photons=args.photons)
    name, cus = tiny_mc_amd.device_info()

    progress(f"device {name} ({cus} CUs); plain walk: {args.warmup} warmup + {args.steps} timed steps")
    wt = tiny_mc_amd.WEIGHT_TABLE_OFF if args.weight_table < 0 else args.weight_table
    for i in range(args.warmup):
        tiny_mc_amd.run_raw(params, 1000 + i, sampler=sampler, weight_table=wt)
    kernel_ms, interactions = [], 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        _, st = tiny_mc_amd.run_raw(params, i + 1, sampler=sampler, weight_table=wt)
        kernel_ms.append(st.kernel_ms)
        interactions += st.interactions
    elapsed = time.perf_counter() - t0  # every call is synchronous
    progress(f"plain walk: {args.steps * args.photons / elapsed:.4g} photons/s")

    photons = args.steps * args.photons
    avg_kernel_s = sum(kernel_ms) / len(kernel_ms) * 1e-3
    flop_per_launch = FLOP_PER_INTERACTION * interactions / args.steps
    achieved = flop_per_launch / avg_kernel_s / 1e12
    kernel_flop = FLOP_PER_INTERACTION_ANALYTIC if args.sampler == "analytic" else FLOP_PER_INTERACTION
    result = {
        "metric": METRIC,
        "value": photons / elapsed,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain walk at scale: {args.photons} photons, default physics (mu_a=2/cm, mu_s=20/cm, "
                        "101 shells x 50 um, cutoff 0.001, roulette 0.1), one launch per step",
            "photons_per_step": args.photons,
            "shells": 101,
            "sampler": args.sampler,
            "transport": "fp32",
            "tallies": "int64 fixed point (128-bit global sums)",
            "device": name,
            "compute_units": cus,
        },
        "interactions_per_s": interactions / elapsed,
        "interactions_per_photon": interactions / photons,
        "kernel_ms_avg": avg_kernel_s * 1e3,
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": read_traffic(),
            "kernel": "walk_f32",
            "flop_per_launch": flop_per_launch,
            "flop_per_absorption": FLOP_PER_INTERACTION,
        },
        "flop_per_absorption_kernel": {"analytic": FLOP_PER_INTERACTION_ANALYTIC, "rejection": FLOP_PER_INTERACTION},
        "frac_kernel_flops": kernel_flop * interactions / args.steps / avg_kernel_s / 1e12 / PEAK_FP32_TFLOPS,
        "weight_table_entries": st.weight_table,
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(),
    }
    if result["cpu_baseline"]:
        progress(f"cpu baseline: {result['cpu_baseline']['value']:.4g} photons/s")
    if not args.no_rejection and args.sampler == "analytic":
        result["rejection"] = rejection_leg(params, min(args.steps, 5))
        progress(f"rejection sampler: {result['rejection']['photons_per_s']:.4g} photons/s")
    if not args.no_reference_case:
        result["reference_case"] = reference_case_leg(sampler)
    if not args.no_sweep:
        result["sweep"] = sweep_leg()
        progress(f"sweep: {result['sweep']['photons_per_s']:.4g} photons/s, "
                 f"{result['sweep']['speedup_vs_one_launch_per_set']:.3f}x one launch per set")
    if not args.no_serve:
        cap = result["sweep"]["interactions_per_s"] if "sweep" in result else interactions / elapsed
        cap = args.serve_capacity or cap
        mcs = args.serve_max_chunk_scale
        result["serve"], _, _ = serve_leg(cap, load=args.serve_load, seconds=args.serve_seconds, max_chunk_scale=mcs)
        progress(f"serve: p50 {result['serve']['p50_ms']:.3g} ms, p99 {result['serve']['p99_ms']:.3g} ms")
    if not args.no_serve and not args.no_saturation:
        sat, reqs, _ = serve_leg(cap, load=1.0, seconds=args.saturation_seconds, seed=13, max_chunk_scale=mcs)
        progress(f"serve at saturation: {sat['interactions_per_s']:.4g} absorptions/s, p50 {sat['p50_ms']:.3g} ms, "
                 f"p99 {sat['p99_ms']:.3g} ms; now one launch per request")
        plain = one_launch_per_request_leg(reqs)
        sat["one_launch_per_request"] = plain
        sat["served_vs_one_launch_per_request"] = sat["interactions_per_s"] / plain["interactions_per_s"]
        sat["capacity_vs_batch"] = sat["interactions_per_s"] / cap
        progress("serve: the same requests as one batch")
        same = same_work_batch_leg(reqs)
        sat["same_work_batch"] = same
        sat["capacity_vs_same_work_batch"] = sat["interactions_per_s"] / same["interactions_per_s"]
        result["serve_saturation"] = sat
        progress(f"serve: {sat['served_vs_one_launch_per_request']:.2f}x one launch per request, "
                 f"{sat['capacity_vs_same_work_batch']:.3f} of the same work batched")
    print(json.dumps(result))


if __name__ == "__main__":
    main()
