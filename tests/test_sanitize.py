"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

tests/sanitize/Makefile builds, with gcc/g++ -fsanitize=address,undefined
-fno-sanitize-recover=all:
- plan_check: the library's host planning (tiny_mc_amd/csrc/tmc_plan.cpp,
  the same source the library is built from) over edge cases and seeded
  random inputs, with the invariants the kernels rely on;
- oracle_check: every entry point of oracle/oracle.c on edge-case inputs;
- tiny_mc_san: the C driver (tiny_mc_amd/host/tiny_mc.c) linked against the
  FP64 oracle in place of libtiny_mc_gpu.so, so its table and error paths run
  without a device;
- tiny_mc_san_rej: the same driver built with -DSAMPLER=1, whose call goes
  through tiny_mc_gpu_run_ex and tiny_mc_gpu_tally_to_heat (the FP32 mirror
  with the rejection sampler behind them).
Any sanitizer report ends the binary with a non-zero status, and LeakSanitizer
checks for leaks at exit.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import oracle_lib

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")
BUILD = os.path.join(HERE, "_build")

SAN_ENV = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


@pytest.fixture(scope="module")
def built():
    if shutil.which("gcc") is None or shutil.which("g++") is None:
        pytest.skip("gcc/g++ not available")
    subprocess.run(["make", "-s", "-C", HERE, "-j3"], check=True, capture_output=True, timeout=300)
    return BUILD


def run(exe, env=None, check=True):
    r = subprocess.run([os.path.join(BUILD, exe)], capture_output=True, text=True, timeout=300,
                       env=env or SAN_ENV)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    if check:
        assert r.returncode == 0, r.stderr[-4000:]
    return r


def test_host_planning_under_sanitizers(built):
    assert run("plan_check").stdout.strip() == "plan_check: ok"


def test_oracle_under_sanitizers(built):
    assert run("oracle_check").stdout.strip() == "oracle_check: ok"


def test_driver_under_sanitizers(built, tmc, orc):
    """The driver's table (tiny_mc.c:37-66 format) from the FP64 oracle's
    heat for SEED=12345, PHOTONS=4096, default physics."""
    lines = run("tiny_mc_san").stdout.splitlines()
    assert lines[5] == "# Photons    =     4096"
    assert re.fullmatch(r"# [0-9]+\.[0-9]{6} seconds", lines[7])
    rows = [ln.split("\t") for ln in lines[11:-1]]
    assert len(rows) == 100 and lines[-1].startswith("# extra\t")
    got = np.array([[float(v) for v in r] for r in rows])
    heat, heat2, _ = orc.walk64(orc.params(), 12345, 4096, nthreads=2)
    want, extra = tmc.table_rows(heat, heat2, 4096, 50.0)
    assert np.all(np.abs(got - np.array(want)) <= 1.5e-5 + 1e-9 * np.abs(np.array(want)))
    assert abs(float(lines[-1].split("\t")[1]) - extra) <= 1.5e-5


def test_rejection_driver_under_sanitizers(built, tmc, orc):
    """SAMPLER=1 build: the table from the FP32 mirror's rejection-sampler
    tallies (own math) for SEED=12345, PHOTONS=4096, default physics."""
    lines = run("tiny_mc_san_rej").stdout.splitlines()
    assert lines[5] == "# Photons    =     4096"
    rows = [ln.split("\t") for ln in lines[11:-1]]
    assert len(rows) == 100 and lines[-1].startswith("# extra\t")
    got = np.array([[float(v) for v in r] for r in rows])
    tally, _ = orc.mirror32(orc.params(), 12345, 4096, sampler=oracle_lib.SAMPLER_REJECTION, nthreads=2)
    heat, heat2 = oracle_lib.tally_heat(tally, 101, 2.0, 20.0)
    want, extra = tmc.table_rows(heat, heat2, 4096, 50.0)
    assert np.all(np.abs(got - np.array(want)) <= 1.5e-5 + 1e-9 * np.abs(np.array(want)))
    assert abs(float(lines[-1].split("\t")[1]) - extra) <= 1.5e-5
    # a different sampler gives a different table for the same stream
    heat_a, heat2_a = oracle_lib.tally_heat(orc.mirror32(orc.params(), 12345, 4096, nthreads=2)[0], 101, 2.0, 20.0)
    assert not np.allclose(tmc.table_rows(heat_a, heat2_a, 4096, 50.0)[0], want)


def test_driver_error_path_under_sanitizers(built):
    for exe in ("tiny_mc_san", "tiny_mc_san_rej"):
        r = run(exe, env=dict(SAN_ENV, TMC_FAKE_STATUS="-2"), check=False)
        assert r.returncode == 1
        assert r.stderr.strip() == "tiny_mc: no device"
