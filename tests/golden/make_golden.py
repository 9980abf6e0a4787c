"""Writes the golden fixtures of tests/golden/ from the UNMODIFIED reference.

Run in a container that holds the reference sources (see oracle/Makefile,
target `ref`, which compiles them where they lie into oracle/_ref/):

    make -C oracle ref && python tests/golden/make_golden.py

Fixtures (all plain text, small):
  ref_<variant>_s<seed>_n<photons>.txt
      the reference's stdout without its two timing lines
      (tiny_mc.c:55-56), for exact comparison with the oracle's restatement;
  stats_<variant>_<seeds>x<photons>.csv   (variants *_sm: rand() replaced
      at link time by SplitMix64, see oracle/ref_rand_splitmix.c)
      per shell: mean and sample standard deviation over seeds 1..<seeds> of
      the reference's heat column (tiny_mc.c:63), and of '# extra'
      (tiny_mc.c:66), for statistical parity of other random streams;
  MANIFEST.json
      compiler, glibc, build flags and sha256 of every fixture.
The fixtures hold only numbers printed by the reference, never its source.
"""
import hashlib
import json
import os
import platform
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")

VARIANTS = {
    # name: (binary, physics recorded for the tests)
    "default": ("tiny_mc_default", dict(mu_a=2.0, mu_s=20.0, shells=101, microns_per_shell=50.0)),
    "lowalb": ("tiny_mc_lowalb", dict(mu_a=10.0, mu_s=10.0, shells=51, microns_per_shell=100.0)),
    "highalb": ("tiny_mc_highalb", dict(mu_a=0.5, mu_s=50.0, shells=201, microns_per_shell=20.0)),
    # the same unmodified sources with rand()/srand() replaced at link time by
    # SplitMix64 (oracle/ref_rand_splitmix.c): the reference's physics
    # without glibc rand()'s lag correlations
    "default_sm": ("tiny_mc_default_sm", dict(mu_a=2.0, mu_s=20.0, shells=101, microns_per_shell=50.0)),
    "highalb_sm": ("tiny_mc_highalb_sm", dict(mu_a=0.5, mu_s=50.0, shells=201, microns_per_shell=20.0)),
}

TABLES = [
    ("default", 12345, 32768),
    ("default", 1, 32768),
    ("default", 7, 1000),
    ("default", 2024, 1),
    ("lowalb", 3, 16384),
    ("highalb", 11, 2048),
]

STATS = [("default", 256, 32768), ("lowalb", 128, 32768), ("highalb", 128, 4096),
         # many short runs: high precision while the reference's float
         # accumulation (tiny_mc.c:26-27) stays well below the noise
         ("default_sm", 8192, 8192), ("highalb_sm", 8192, 1024)]

TIMING = re.compile(r"^# [0-9.]+ (seconds|K photons per second)$")


def run_ref(variant, seed, photons):
    exe = os.path.join(REF, VARIANTS[variant][0])
    env = dict(os.environ, TMC_SEED=str(seed), TMC_PHOTONS=str(photons))
    out = subprocess.run([exe], env=env, check=True, capture_output=True, text=True).stdout
    return "".join(line + "\n" for line in out.splitlines() if not TIMING.match(line))


def parse_table(text):
    heat, extra = [], None
    for line in text.splitlines():
        if line.startswith("# extra"):
            extra = float(line.split("\t")[1])
        elif not line.startswith("#"):
            heat.append(float(line.split("\t")[1]))
    return np.array(heat), extra


def sha(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def main():
    for v in VARIANTS.values():
        if not os.path.exists(os.path.join(REF, v[0])):
            sys.exit(f"missing {REF}/{v[0]}: run make -C oracle ref first")
    manifest = {
        "compiler": subprocess.run(["gcc", "--version"], capture_output=True, text=True).stdout.splitlines()[0],
        "glibc": platform.libc_ver()[1],
        "flags": "-std=c11 -O2 -ffp-contract=off -DSEED='atoi(getenv(\"TMC_SEED\"))' "
                 "-DPHOTONS='atoi(getenv(\"TMC_PHOTONS\"))' (+ per-variant -D physics; *_sm variants also "
                 "-Wl,--wrap=rand,--wrap=srand oracle/ref_rand_splitmix.c)",
        "variants": {k: v[1] for k, v in VARIANTS.items()},
        "files": {},
    }
    for variant, seed, photons in TABLES:
        name = f"ref_{variant}_s{seed}_n{photons}.txt"
        path = os.path.join(HERE, name)
        with open(path, "w") as f:
            f.write(run_ref(variant, seed, photons))
        manifest["files"][name] = {"variant": variant, "seed": seed, "photons": photons, "sha256": sha(path)}
    workers = min(8, os.cpu_count() or 1)
    for variant, nseeds, photons in STATS:
        with ThreadPoolExecutor(workers) as ex:
            tables = list(ex.map(lambda s: parse_table(run_ref(variant, s, photons)), range(1, nseeds + 1)))
        heat = np.array([t[0] for t in tables])
        extra = np.array([t[1] for t in tables])
        name = f"stats_{variant}_{nseeds}x{photons}.csv"
        path = os.path.join(HERE, name)
        with open(path, "w") as f:
            f.write("# shell,mean_heat,sd_heat  (reference heat column, W/cm^3; seeds 1..%d)\n" % nseeds)
            for i in range(heat.shape[1]):
                f.write("%d,%.9g,%.9g\n" % (i, heat[:, i].mean(), heat[:, i].std(ddof=1)))
            f.write("extra,%.9g,%.9g\n" % (extra.mean(), extra.std(ddof=1)))
        manifest["files"][name] = {"variant": variant, "seeds": nseeds, "photons": photons, "sha256": sha(path)}
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
