"""The headless viewer's frames (cg_mc.c:71-87: 20 photons per frame) on the
GPU path: after every frame the storage buffer equals the heat of one run of
all the photons so far, bit for bit (integer tallies add exactly)."""
import numpy as np
import pytest

import tiny_mc_amd as tmc
from tiny_mc_amd import viewer

pytestmark = pytest.mark.gpu


def test_frames_equal_one_run_of_the_same_photons():
    p = tmc.Params()
    seen = 0
    for done, heats in viewer.frames(p, 11, photons_per_frame=20, cap=1000):
        seen += 1
        if done in (20, 500, 1000):
            run = tmc.Params(p.mu_a, p.mu_s, p.shells, p.microns_per_shell, done)
            tally, _ = tmc.run_raw(run, 11)
            want = np.asarray(tmc.tally_to_heat(run, tally)[0], dtype=np.float32)
            np.testing.assert_array_equal(heats, want)
    assert seen == 50
    red = viewer.render(heats, size=64)
    assert red.max() > 0.0
