"""Headless counterpart of the reference's OpenGL viewer (cg_mc.c).

The viewer walks MAX_PHOTONS_PER_FRAME = 20 photons per frame into a float
``heats[SHELLS]`` array until PHOTON_CAP = 2^16 photons are done
(cg_mc.c:71-87), uploads the array as a shader storage buffer, and its
fragment shader colours each pixel of an 800 x 800 window by the heat of the
shell at the pixel's distance from the centre (cg_mc.c:37-66).  Here:

- ``frames`` yields the storage buffer's contents after each frame: the
  frame's photons are one library call (photon indices frame * 20 onwards
  under one seed), the integer tallies add exactly across calls, and the
  heat is converted once per frame, as ``tiny_mc_gpu_tally_to_heat`` does,
  then rounded to float32 (the buffer's type).
- ``render`` evaluates the fragment shader on the CPU for every pixel and
  returns the red channel, and ``write_ppm`` saves it as a binary PPM.

``python -m tiny_mc_amd.viewer --out heat.ppm`` runs every frame and writes
the last image.  OpenGL, GLFW and GLEW are not needed.
"""
import argparse
import math

import numpy as np

from . import _lib

PHOTON_CAP = 1 << 16  # cg_mc.c:10
MAX_PHOTONS_PER_FRAME = 20  # cg_mc.c:11
WINDOW = 800  # cg_mc.c:93, 132

# fragment shader constants (cg_mc.c:41, 58-60)
MC = 0.7071067811865476
L_FIT = 2.0
B_FIT = 1.0
K_FIT = 0.004


def frames(params: _lib.Params, seed: int, photons_per_frame=MAX_PHOTONS_PER_FRAME, cap=PHOTON_CAP):
    """The float32 heats buffer after each frame, as (photons done, heats).

    ``params.photons`` is ignored: the viewer's budget is ``cap``."""
    if photons_per_frame <= 0 or cap < 0:
        raise ValueError("photons_per_frame must be positive and cap non-negative")
    tally = np.zeros(4 * params.shells, dtype=np.uint64)
    done = 0
    while done < cap:
        n = min(photons_per_frame, cap - done)
        frame = _lib.Params(params.mu_a, params.mu_s, params.shells, params.microns_per_shell, n)
        t, _ = _lib.run_raw(frame, seed, first_photon=done)
        add128(tally, t)
        done += n
        yield done, heats_of(params, tally)


def add128(acc, t):
    """acc += t for raw tallies: (low, high) pairs of 64-bit words, each pair
    one 128-bit sum (low word first), with the carry out of the low word."""
    lo = acc[0::2] + t[0::2]  # wraps modulo 2^64
    carry = (lo < acc[0::2]).astype(np.uint64)
    acc[1::2] += t[1::2] + carry
    acc[0::2] = lo


def heats_of(params: _lib.Params, tally):
    """float32 heats (the storage buffer) from accumulated raw tallies."""
    return np.asarray(_lib.tally_to_heat(params, tally)[0], dtype=np.float32)


def render(heats, size=WINDOW):
    """The fragment shader of cg_mc.c:37-66 for every pixel of a size x size
    window: red channel in [0, 1) (green and blue are 0), row 0 at the top.

    gl_FragCoord is the pixel centre (x + 0.5, y + 0.5); the shell index is
    int(dr / MC * (len - 1)) for dr the distance of uv from (0.5, 0.5), and
    the colour L / (1 + b exp(-k heat)) - 1, evaluated in float32 as the
    shader does."""
    heats = np.asarray(heats, dtype=np.float32)
    n = len(heats)
    c = (np.arange(size, dtype=np.float32) + np.float32(0.5)) / np.float32(size)
    u, v = np.meshgrid(c, c[::-1])  # window y grows upwards
    dr = np.sqrt((u - np.float32(0.5)) ** 2 + (v - np.float32(0.5)) ** 2).astype(np.float32)
    idx = ((dr / np.float32(MC)) * np.float32(n - 1)).astype(np.int64)
    heat = heats[np.clip(idx, 0, n - 1)]
    with np.errstate(over="ignore"):
        fit = np.float32(L_FIT) / (np.float32(1.0) + np.float32(B_FIT) * np.exp(np.float32(-K_FIT) * heat)) - np.float32(1.0)
    return fit.astype(np.float32)


def write_ppm(path, red):
    """Binary PPM (P6) of a red channel in [0, 1]."""
    h, w = red.shape
    rgb = np.zeros((h, w, 3), dtype=np.uint8)
    rgb[..., 0] = np.clip(np.rint(red * 255.0), 0, 255).astype(np.uint8)
    with open(path, "wb") as f:
        f.write(b"P6\n%d %d\n255\n" % (w, h))
        f.write(rgb.tobytes())


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--photons", type=int, default=PHOTON_CAP)
    ap.add_argument("--per-frame", type=int, default=MAX_PHOTONS_PER_FRAME)
    ap.add_argument("--out", default="tiny_mc_heat.ppm")
    a = ap.parse_args(argv)
    p = _lib.Params()
    last = None
    for done, heats in frames(p, a.seed, a.per_frame, a.photons):
        last = heats
    if last is None:
        last = np.zeros(p.shells, dtype=np.float32)
    write_ppm(a.out, render(last))
    print(f"{a.out}: {a.photons} photons, {math.ceil(a.photons / a.per_frame)} frames")


if __name__ == "__main__":
    main()
