"""High-albedo precision study of the walk's arithmetic.

At mu_a = 0.01/cm, mu_s = 100/cm (albedo 0.9999, about 7.2e4 absorptions
per photon), 401 shells x 10 um, the reference's binary32 arithmetic
(photon.c:21-49) drifts: `weight *= albedo` (photon.c:32) loses about 1.8e-3
by the cutoff, so the roulette cadence moves (cutoff at absorption 69063
instead of 69082), and positions summed over 7e4 steps carry errors of the
order of the float spacing of |x|.  This module walks the same Philox stream
with every precision variant of tiny_mc_gpu_run_ex and compares each with
the all-FP64 GPU walk:

    python -m tiny_mc_amd.precision --photons 268435456 --parts 16

prints one JSON object (per variant: photons/s, absorptions per photon, the
absorbed energy and the per-shell deviation from FP64, in relative terms and
in units of the FP64 run's per-shell standard error).
"""
import argparse
import json
import sys
import time

import numpy as np

from . import _lib as L

STRESS = dict(mu_a=0.01, mu_s=100.0, shells=401, microns_per_shell=10.0)
MODES = [
    ("fp64", L.PRECISION_FP64, L.MATH_HARDWARE),
    ("fp32", L.PRECISION_FP32, L.MATH_HARDWARE),
    ("fp32_own_math", L.PRECISION_FP32, L.MATH_OWN),
    ("fp32_weight", L.PRECISION_FP32_WEIGHT, L.MATH_HARDWARE),
    ("fp32_compensated", L.PRECISION_FP32_COMPENSATED, L.MATH_HARDWARE),
]


def _add_tallies(acc, tally):
    """128-bit (lo, hi) word pairs added exactly (Python integers)."""
    t = [int(v) for v in tally]
    vals = [t[2 * i] + (t[2 * i + 1] << 64) for i in range(len(t) // 2)]
    return vals if acc is None else [a + b for a, b in zip(acc, vals)]


def _to_words(vals):
    out = np.zeros(2 * len(vals), dtype=np.uint64)
    for i, v in enumerate(vals):
        out[2 * i] = v & ((1 << 64) - 1)
        out[2 * i + 1] = v >> 64
    return out


def run_mode(params, seed, precision, parts=1, progress=None, math=L.MATH_HARDWARE):
    """Walks params.photons photons in `parts` launches over consecutive
    photon ranges (their integer tallies add exactly)."""
    acc, inter, ms = None, 0, 0.0
    N = params.photons
    for i in range(parts):
        lo, hi = N * i // parts, N * (i + 1) // parts
        if hi == lo:
            continue
        p = L.Params(params.mu_a, params.mu_s, params.shells, params.microns_per_shell, hi - lo)
        tally, st = L.run_raw(p, seed, first_photon=lo, precision=precision, math=math)
        acc = _add_tallies(acc, tally)
        inter += st.interactions
        ms += st.kernel_ms
        if progress:
            progress(f"precision {precision}: part {i + 1}/{parts} done, {ms / 1e3:.1f} s of kernel time")
    heat, heat2 = L.tally_to_heat(params, _to_words(acc))
    return heat, heat2, inter, ms


def compare(heat, ref, ref2, photons):
    """Per-shell deviation of `heat` from the FP64 run (shells 0..S-2; the
    last shell is the overflow bin)."""
    h, r, r2 = heat[:-1], ref[:-1], ref2[:-1]
    sigma = np.sqrt(np.maximum(r2 - r * r / photons, 0.0))
    ok = r > 0
    rel = np.where(ok, (h - r) / np.where(ok, r, 1.0), 0.0)
    z = np.where(sigma > 0, (h - r) / np.where(sigma > 0, sigma, 1.0), 0.0)
    w = r / r.sum()
    return {
        "max_abs_rel": float(np.abs(rel).max()),
        "weighted_mean_abs_rel": float((np.abs(rel) * w).sum()),
        "max_abs_z": float(np.abs(z).max()),
        "rms_z": float(np.sqrt((z * z).mean())),
        "shells_beyond_3sigma": int((np.abs(z) > 3.0).sum()),
    }


def study(photons, seed=1, parts=1, modes=MODES, progress=None):
    params = L.Params(photons=photons, **STRESS)
    out = {"physics": dict(STRESS), "photons": photons, "seed": seed, "variants": {}}
    ref = None
    for name, prec, math in modes:
        t0 = time.perf_counter()
        heat, heat2, inter, ms = run_mode(params, seed, prec, parts, progress, math)
        rec = {
            "precision": prec,
            "math": math,
            "kernel_s": ms / 1e3,
            "wall_s": time.perf_counter() - t0,
            "photons_per_s": photons / (ms / 1e3) if ms > 0 else None,
            "interactions_per_photon": inter / photons,
            "energy_per_photon": float(heat.sum() / photons),
        }
        if ref is None:
            ref = (heat, heat2, rec["energy_per_photon"], inter)
        else:
            rec["energy_rel_vs_fp64"] = rec["energy_per_photon"] / ref[2] - 1.0
            rec["interactions_rel_vs_fp64"] = inter / ref[3] - 1.0
            rec["shells_vs_fp64"] = compare(heat, ref[0], ref[1], photons)
        out["variants"][name] = rec
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--photons", type=int, default=1 << 22)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--parts", type=int, default=1, help="launches per variant (progress lines between)")
    args = ap.parse_args(argv)

    def progress(msg):
        print(f"[precision {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)

    print(json.dumps(study(args.photons, args.seed, args.parts, progress=progress)), flush=True)


if __name__ == "__main__":
    main()
