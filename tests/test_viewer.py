"""Headless counterpart of the reference's viewer (cg_mc.c): the shader's
pixel-to-shell mapping and colour, and the exact accumulation of frame
tallies (CPU only)."""
import math

import numpy as np

from tiny_mc_amd import viewer


def test_add128_carries_out_of_the_low_word():
    acc = np.array([2 ** 64 - 5, 7, 1, 0], dtype=np.uint64)
    viewer.add128(acc, np.array([10, 2, 3, 4], dtype=np.uint64))
    assert acc.tolist() == [5, 10, 4, 4]  # 7 + 2 + carry
    # planted mistake: a plain word-wise sum loses the carry
    plain = np.array([2 ** 64 - 5, 7], dtype=np.uint64) + np.array([10, 2], dtype=np.uint64)
    assert plain.tolist() == [5, 9]


def test_shader_maps_pixels_to_shells_by_distance():
    n = 101
    heats = np.arange(n, dtype=np.float32) * 10.0  # heat 10 i in shell i
    red = viewer.render(heats)
    assert red.shape == (800, 800)
    # the inverse of the colour map: heat = -ln(L / (red + 1) - 1) / k (b = 1)
    def shell_of(r):
        return int(round(-math.log(2.0 / (float(r) + 1.0) - 1.0) / 0.004 / 10.0)) if r < 0.999 else None
    # the four centre pixels lie sqrt(2) * 0.5 / 800 from the centre: shell 0
    for y, x in ((399, 399), (399, 400), (400, 399), (400, 400)):
        assert red[y, x] == 0.0
    # pixel (row 400, column 799): centre (799.5, 399.5) in window coordinates
    # (row 400 from the top is y = 399 from the bottom), distance 399.5 / 800
    # from the centre: shell int(0.499375 / MC * 100) = 70
    assert shell_of(red[400, 799]) == int(0.499375 / viewer.MC * 100) == 70
    # pixel (row 0, column 0): distance sqrt(2) * 399.5 / 800: shell 99
    assert shell_of(red[0, 0]) == int(math.sqrt(2) * 399.5 / 800 / viewer.MC * 100) == 99
    # radially non-decreasing and symmetric under the window's reflections
    row = red[400, 400:]
    assert np.all(np.diff(row) >= 0)
    np.testing.assert_array_equal(red, red[:, ::-1])
    np.testing.assert_array_equal(red, red[::-1, :])
    np.testing.assert_array_equal(red, red.T)


def test_shader_colour_is_black_at_zero_heat_and_saturates():
    red = viewer.render(np.zeros(101, dtype=np.float32), size=16)
    assert np.all(red == 0.0)
    red = viewer.render(np.full(101, 1e9, dtype=np.float32), size=16)
    assert np.all(red == 1.0)


def test_ppm_header_and_size(tmp_path):
    red = viewer.render(np.linspace(0, 2000, 51, dtype=np.float32), size=32)
    path = tmp_path / "h.ppm"
    viewer.write_ppm(path, red)
    data = path.read_bytes()
    assert data.startswith(b"P6\n32 32\n255\n")
    assert len(data) == len(b"P6\n32 32\n255\n") + 32 * 32 * 3
