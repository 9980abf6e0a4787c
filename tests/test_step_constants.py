"""Constants of the walk's step that replace a floating-point comparison by
an integer one (CPU only).

The roulette (photon.c:46) kills a photon when u > 0.1f. The kernel's
uniform is u = ((r >> 8) + 1) 2^-24 of a 32-bit Philox word r, so the test
is a threshold on r; the threshold in tmc_step.h is checked here against
the float comparison on every one of the 2^24 uniforms.

Whether an absorption's weight w falls below the cutoff (photon.c:45),
RN(w a) < 0.001f, is read as fx_heat(w) < h_cut = fx_heat(T), T the least
float with RN(T a) >= 0.001f (tmc_step.h, WeightTable); checked here against
the float comparison over the binary32 weight sequences of many albedos, on
random weights and on every float near T.
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


def _weight_sequence(albedo, n):
    """W[0 .. n-1] of the binary32 walk (photon.c:32, 45-49): W[0] = 1,
    W[k+1] = RN(W[k] a), divided by 0.1f when below the cutoff (IEEE binary32
    multiply and divide, as the kernel's table builder), and whether each
    RN(W[k] a) falls below the cutoff."""
    a = np.float32(albedo)
    w = np.float32(1.0)
    ws = np.empty(n, np.float32)
    cut = np.empty(n, bool)
    for k in range(n):
        ws[k] = w
        w2 = np.float32(w * a)
        cut[k] = w2 < np.float32(0.001)
        w = np.float32(w2 / np.float32(0.1)) if cut[k] else w2
    return ws, cut


def _threshold(albedo):
    """T by bisection over the bit patterns of (0, 1], as the kernel's table
    builder; None when even 1 * a falls below the cutoff."""
    a = np.float32(albedo)
    if np.float32(np.float32(1.0) * a) < np.float32(0.001):
        return None
    lo, hi = 0, int(np.float32(1.0).view(np.uint32))
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if np.float32(np.uint32(mid).view(np.float32) * a) < np.float32(0.001):
            lo = mid
        else:
            hi = mid
    return np.uint32(hi).view(np.float32)


def _fx_heat(w):
    return np.floor(np.asarray(w, np.float32).astype(np.float64) * 2.0 ** 33).astype(np.uint64)


def test_cut_is_a_threshold_on_the_weight_image():
    rng = np.random.default_rng(7)
    # albedos from 0.001 to 0.9999 (log-spaced), the default 20/22 and round values
    albedos = list(1.0 - np.geomspace(0.999, 1e-4, 160)) + [20.0 / 22.0, 0.05, 0.1, 0.5, 0.0005]
    checked = 0
    for a in albedos:
        a32 = np.float32(a)
        if a32 >= np.float32(1.0):
            continue
        T = _threshold(a)
        h_cut = np.uint64(2 ** 64 - 1) if T is None else _fx_heat(T)
        # the weight sequence, random weights in [0, 1] and the 2^16 floats around T
        ws, cut = _weight_sequence(a, 1024)
        rand = rng.random(20000).astype(np.float32)
        tiny = (rng.random(2000) * 2.0 ** -12).astype(np.float32)
        near = []
        if T is not None:
            tb = int(np.float32(T).view(np.uint32))
            near = np.arange(tb - 2 ** 15, tb + 2 ** 15, dtype=np.int64).astype(np.uint32).view(np.float32)
        w = np.concatenate([ws, rand, tiny, np.asarray(near, np.float32), np.float32([0.0, 1.0])])
        w = w[(w >= 0) & (w <= 1)]
        truth = (w * a32).astype(np.float32) < np.float32(0.001)
        np.testing.assert_array_equal(_fx_heat(w) < h_cut, truth, err_msg=f"albedo {a}")
        np.testing.assert_array_equal(_fx_heat(ws) < h_cut, cut, err_msg=f"albedo {a}")
        if T is not None:
            assert T > np.float32(2.0 ** -10)  # where fx_heat is exact
        checked += 1
    assert checked > 160
    # planted mistakes: a threshold one image off either way misreads T or
    # the float below it at the default albedo
    a32 = np.float32(20.0 / 22.0)
    T = _threshold(20.0 / 22.0)
    tb = int(np.float32(T).view(np.uint32))
    w = np.arange(tb - 4, tb + 4, dtype=np.uint32).view(np.float32)
    truth = (w * a32).astype(np.float32) < np.float32(0.001)
    np.testing.assert_array_equal(_fx_heat(w) < _fx_heat(T), truth)
    for h_bad in (_fx_heat(T) + np.uint64(1), _fx_heat(T) - np.uint64(1)):
        assert not np.array_equal(_fx_heat(w) < h_bad, truth)
