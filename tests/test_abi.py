"""The C ABI library (CPU-side checks; no device work).

- libtiny_mc_gpu.so loads and exports every function include/tiny_mc_gpu.h
  declares;
- the header compiles as C and C++;
- parameter validation rejects what the reference cannot run (params.h:14:
  mu_a must be > 0) before any device is touched;
- the host conversion of raw 128-bit tallies to heat / heat2 matches an
  independent computation;
- the driver binaries are linked against the library.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tiny_mc_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*[\s\*]+([a-z_][a-z0-9_]*)\s*\(", text, re.M)
    return sorted(set(names))


def test_every_declared_symbol_is_exported(tmc):
    names = declared_functions()
    assert "tiny_mc_gpu_run" in names and "tiny_mc_gpu_run_ex" in names
    lib = ctypes.CDLL(tmc.lib_path)
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", tmc.lib_path], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(names) <= exported
    # nothing else leaks out of the library but the ABI
    assert {n for n in exported if not n.startswith("_")} <= set(names)


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_header_compiles(tmp_path, lang):
    src = tmp_path / ("t.c" if lang == "c" else "t.cpp")
    src.write_text('#include "tiny_mc_gpu.h"\nint main(void){ tmc_params p = {2.0f, 20.0f, 101, 50.0f, 1};'
                   ' (void)p; return TMC_OK; }\n')
    cc = "gcc" if lang == "c" else "g++"
    std = "-std=c99" if lang == "c" else "-std=c++11"
    subprocess.run([cc, std, "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.dirname(HEADER), str(src)],
                   check=True)


def test_invalid_params_rejected_without_device(tmc):
    for bad in (dict(mu_a=0.0), dict(mu_a=-1.0), dict(mu_s=-0.5), dict(shells=0), dict(shells=tmc.MAX_SHELLS + 1),
                dict(microns_per_shell=0.0), dict(mu_a=float("nan")), dict(mu_s=float("inf")),
                dict(mu_a=1e-9, mu_s=100.0)):
        p = tmc.Params(**bad)
        with pytest.raises(tmc.TmcError) as e:
            tmc.run_raw(p, 1)
        assert e.value.status == -1
        with pytest.raises(tmc.TmcError) as e:
            tmc.run(p, 1)
        assert e.value.status == -1


def test_strerror_and_defaults(tmc):
    lib = tmc.load()
    assert lib.tiny_mc_gpu_strerror(0) == b"ok"
    assert lib.tiny_mc_gpu_strerror(-1) == b"invalid argument"
    assert lib.tiny_mc_gpu_strerror(12345) == b"unknown status"
    from tiny_mc_amd._lib import _Options

    o = _Options(7, 7, 7, 7, 7, 7, 7, 7, 7, 7)
    lib.tmc_options_default(ctypes.byref(o))
    assert (o.sampler, o.block, o.grid, o.chunk, o.first_photon, o.max_interactions, o.precision,
            o.math, o.weight_table, o.reserved) == (0, 0, 0, 0, 0, 0, 0, 0, 0, 0)
    assert ctypes.sizeof(o) == 48


def test_tally_to_heat_matches_independent_conversion(tmc):
    rng = np.random.default_rng(3)
    for phys in (dict(), dict(mu_a=0.01, mu_s=100.0, shells=401, microns_per_shell=10.0)):
        p = tmc.Params(**phys)
        S = p.shells
        t = rng.integers(0, 2**63, size=4 * S, dtype=np.uint64)
        t[1::2] = rng.integers(0, 4, size=2 * S, dtype=np.uint64)  # high words small
        h, h2 = tmc.tally_to_heat(p, t)
        eh, eh2 = oracle_lib.tally_heat(t, S, p.mu_a, p.mu_s)
        np.testing.assert_allclose(h, eh, rtol=1e-15)
        np.testing.assert_allclose(h2, eh2, rtol=1e-15)


def test_driver_binaries_link_the_library(tmc):
    for name in ("tiny_mc", "tiny_mc_seed12345", "tiny_mc_seed12345_rej"):
        exe = os.path.join(ROOT, "tiny_mc_amd", name)
        assert os.path.exists(exe), exe
        out = subprocess.run(["ldd", exe], capture_output=True, text=True, check=True).stdout
        assert "libtiny_mc_gpu.so" in out and "not found" not in out
