"""Constants of the walk's step that replace a floating-point comparison by
an integer one (CPU only).

The roulette (photon.c:46) kills a photon when u > 0.1f. The kernel's
uniform is u = ((r >> 8) + 1) 2^-24 of a 32-bit Philox word r, so the test
is a threshold on r; the threshold in tmc_step.h is checked here against
the float comparison on every one of the 2^24 uniforms.
"""
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
STEP_H = os.path.join(HERE, "..", "tiny_mc_amd", "csrc", "tmc_step.h")


def _define(name):
    text = open(STEP_H).read()
    m = re.search(r"#define\s+" + name + r"\s+(0x[0-9a-fA-F]+)u", text)
    assert m, name
    return int(m.group(1), 16)


def test_roulette_kill_threshold_equals_float_compare():
    kill_r = _define("TMC_ROULETTE_KILL_R")
    m = np.arange(1 << 24, dtype=np.uint64)
    u = ((m + 1).astype(np.float64) * 2.0 ** -24).astype(np.float32)  # exact
    kills = u > np.float32(0.1)
    # r >= kill_r depends on r >> 8 only when the threshold's low byte is 0
    assert kill_r & 0xFF == 0
    np.testing.assert_array_equal(kills, (m << np.uint64(8)) >= np.uint64(kill_r))
    # about 90% of the draws kill (survival probability 0.1f)
    assert abs(kills.mean() - 0.9) < 1e-7
