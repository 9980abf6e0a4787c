"""Benchmark of the MI355X photon-transport path (one GPU).

A step is one launch of the walk over 2^30 photons at default physics
(mu_a = 2/cm, mu_s = 20/cm, 101 shells of 50 um, weight cutoff 0.001,
roulette 0.1), through the C ABI call tiny_mc_gpu_run_ex: clear the tallies,
run the persistent kernel, copy the 128-bit tallies back.  There is no input
data: photons are generated from (seed, photon index) on the device.

Prints one JSON line: whole-job photons/s, the roofline of the walk kernel
(algorithmic FP32 FLOPs of the reference algorithm, 40.5 per absorption
event counted from photon.c:21-48, over the kernel's average duration
measured with HIP events on the library's stream), and the reference CPU
loop timed on this machine's host cores.

Two more legs of the metric follow, each once:
  sweep   4096 parameter sets (mu_a log-uniform [0.05,10]/cm, mu_s
          log-uniform [10,200]/cm, shells {51,101,201,401}, width U[10,100] um,
          photons log-uniform [2^16,2^22], seed = set index) in one persistent
          launch, beside the same sets run one launch per set;
  serve   Poisson arrivals of requests (sweep physics, Pareto(1.5) photons in
          [2^14,2^24]) at a fixed fraction of the sweep's measured capacity;
          p50/p99 latency, aggregate photons/s, and a sample of results
          checked bit for bit against the same request run alone.

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
FLOP_PER_INTERACTION = 40.5  # FP32 FLOPs of photon.c:21-48 per absorption event
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


def cpu_baseline(max_procs=16, photons_per_proc=1 << 19):
    """The reference loop (photon() + glibc rand(), unmodified sources built
    by oracle/Makefile) on the host cores: one process per core, distinct
    seeds (threads would serialise on rand()'s lock).  Σ photons / wall."""
    procs = max(1, min(max_procs, len(os.sched_getaffinity(0))))
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
    rates = []
    for p in running:
        out, _ = p.communicate()
        if p.returncode != 0:
            return None
        m = re.search(r"^# ([0-9.]+) K photons per second", out, re.M)
        if m:
            rates.append(float(m.group(1)) * 1e3)
    wall = time.perf_counter() - t0
    shipped = ref_rate(REF_SHIPPED, 1, photons_per_proc // 2) if os.path.exists(REF_SHIPPED) else None
    return {
        "value": procs * photons_per_proc / wall,
        "unit": "photons/s",
        "cores": procs,
        "kind": kind,
        "cpu_model": cpu_model(),
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


def sweep_leg():
    """The batched sweep in one launch vs one launch per set (same sets)."""
    import tiny_mc_amd

    sets = tiny_mc_amd.sweep_sets(n=4096, seed=0)
    photons = sum(p.photons for p, _ in sets)
    # warm the multi-set kernel's code object on a small batch
    tiny_mc_amd.batch_run(tiny_mc_amd.sweep_sets(n=8, seed=1, photon_scale=1e-3))
    with tiny_mc_amd.BatchPlan(sets, tiny_mc_amd.BatchOptions(use_graph=True)) as plan:
        t0 = time.perf_counter()
        _, st = plan.run()
        batched_s = time.perf_counter() - t0
    progress(f"sweep batched: {batched_s:.2f} s; now one launch per set")
    t0 = time.perf_counter()
    for p, seed in sets:
        tiny_mc_amd.run_raw(p, seed)
    per_set_s = time.perf_counter() - t0
    return {
        "sets": len(sets),
        "photons": photons,
        "interactions": st.interactions,
        "seconds": batched_s,
        "photons_per_s": photons / batched_s,
        "interactions_per_s": st.interactions / batched_s,
        "one_launch_per_set_seconds": per_set_s,
        "one_launch_per_set_photons_per_s": photons / per_set_s,
        "speedup_vs_one_launch_per_set": per_set_s / batched_s,
        "chunks": st.chunks,
        "lane_utilisation": st.lane_utilisation,
        "launch": {"passes": st.passes, "grid": st.grid, "block": st.block, "waves_per_simd": st.waves_per_simd,
                   "copies": st.copies, "chunk_work": st.chunk_work, "hip_graph": True},
    }


def serve_leg(capacity_interactions_per_s, load=0.5, seconds=15.0, check=16):
    """Open-loop Poisson load at `load` x capacity for about `seconds`."""
    import numpy as np

    import tiny_mc_amd

    probe = tiny_mc_amd.serve_load.draw_requests(4000, seed=11)
    mean_work = tiny_mc_amd.serve_load.mean_request_work(probe)
    rate = load * capacity_interactions_per_s / mean_work
    n = max(50, int(rate * seconds))
    reqs = tiny_mc_amd.serve_load.draw_requests(n, seed=12)
    with tiny_mc_amd.Server() as srv:
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
    }


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
    ap.add_argument("--serve-seconds", type=float, default=15.0)
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
        },
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(),
    }
    if result["cpu_baseline"]:
        progress(f"cpu baseline: {result['cpu_baseline']['value']:.4g} photons/s")
    if not args.no_reference_case:
        result["reference_case"] = reference_case_leg(sampler)
    if not args.no_sweep:
        result["sweep"] = sweep_leg()
        progress(f"sweep: {result['sweep']['photons_per_s']:.4g} photons/s, "
                 f"{result['sweep']['speedup_vs_one_launch_per_set']:.3f}x one launch per set")
    if not args.no_serve:
        cap = result["sweep"]["interactions_per_s"] if "sweep" in result else interactions / elapsed
        result["serve"] = serve_leg(cap, load=args.serve_load, seconds=args.serve_seconds)
        progress(f"serve: p50 {result['serve']['p50_ms']:.3g} ms, p99 {result['serve']['p99_ms']:.3g} ms")
    print(json.dumps(result))


if __name__ == "__main__":
    main()
