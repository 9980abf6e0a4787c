"""The tiny_mc driver on the GPU path prints the reference's table
(tiny_mc.c:37-66 format) with the values the oracle's FP32 mirror gives for
the same stream."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("exe_name,sampler", [("tiny_mc_seed12345", 0), ("tiny_mc_seed12345_rej", 1)])
def test_driver_table_matches_mirror(tmc, orc, exe_name, sampler):
    """SEED=12345, PHOTONS=65536; the _rej build (-DSAMPLER=1) walks with the
    reference's rejection sampler."""
    exe = os.path.join(ROOT, "tiny_mc_amd", exe_name)
    out = subprocess.run([exe], capture_output=True, text=True, check=True, timeout=120).stdout
    lines = out.splitlines()
    assert lines[0] == "# Tiny Monte Carlo by Scott Prahl (http://omlc.ogi.edu)"
    assert lines[1] == "# 1 W Point Source Heating in Infinite Isotropic Scattering Medium"
    assert lines[3] == "# Scattering =   20.000/cm"
    assert lines[4] == "# Absorption =    2.000/cm"
    assert lines[5] == "# Photons    =    65536"
    assert lines[6] == "#"
    assert re.fullmatch(r"# [0-9]+\.[0-9]{6} seconds", lines[7])
    assert re.fullmatch(r"# [0-9]+\.[0-9]{6} K photons per second", lines[8])
    assert lines[9] == "# Radius\tHeat" and lines[10] == "# [microns]\t[W/cm^3]\tError"
    rows = [ln.split("\t") for ln in lines[11:-1]]
    assert len(rows) == 100 and lines[-1].startswith("# extra\t")
    got = np.array([[float(v) for v in r] for r in rows])

    tally, _ = orc.mirror32(orc.params(), 12345, 65536, sampler=sampler,
                            tables=tmc.step_math_tables(tmc.MATH_HARDWARE))
    heat, heat2 = oracle_lib.tally_heat(tally, 101, 2.0, 20.0)
    rows_want, extra = tmc.table_rows(heat, heat2, 65536, 50.0)
    want = np.array(rows_want)
    # printed with 5 decimals (%12.5f)
    assert np.all(np.abs(got - want) <= 1.5e-5 + 1e-9 * np.abs(want))
    assert abs(float(lines[-1].split("\t")[1]) - extra) <= 1.5e-5
    for i, r in enumerate(rows):
        assert r[0] == "%6.0f" % (i * 50.0)
