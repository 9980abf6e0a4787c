"""Philox4x32-10 against the published known-answer vectors (Random123,
Salmon et al. SC'11): the oracle's implementation, the product header
compiled for the host, and (GPU) the device kernel."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_oracle_philox_kat(orc, ctr, key, want):
    assert tuple(orc.philox(ctr, key)) == want


def test_product_header_on_host(orc, tmp_path):
    """tiny_mc_amd/csrc/tmc_philox.h is plain C outside HIP: compile it with
    gcc and check the vectors plus 1000 pseudo-random counters against the
    oracle's implementation."""
    src = tmp_path / "kat.c"
    src.write_text(
        '#include <stdio.h>\n#include "tmc_philox.h"\n'
        "int main(void){ unsigned s = 1;\n"
        "  for (int i = 0; i < 1003; ++i) { unsigned c[6];\n"
        "    for (int j = 0; j < 6; ++j) { s = s * 1664525u + 1013904223u; c[j] = s; }\n"
        "    if (i == 0) { c[0]=c[1]=c[2]=c[3]=c[4]=c[5]=0; }\n"
        "    if (i == 1) { c[0]=c[1]=c[2]=c[3]=c[4]=c[5]=0xFFFFFFFFu; }\n"
        "    if (i == 2) { c[0]=0x243F6A88u; c[1]=0x85A308D3u; c[2]=0x13198A2Eu; c[3]=0x03707344u;"
        " c[4]=0xA4093822u; c[5]=0x299F31D0u; }\n"
        "    tmc_u32x4 r = tmc_philox4x32_10(c[0],c[1],c[2],c[3],c[4],c[5]);\n"
        "    printf(\"%u %u %u %u %u %u %u %u %u %u\\n\", c[0],c[1],c[2],c[3],c[4],c[5],r.x0,r.x1,r.x2,r.x3); }\n"
        "  return 0; }\n")
    exe = tmp_path / "kat"
    subprocess.run(["gcc", "-std=c11", "-O2", "-I", os.path.join(ROOT, "tiny_mc_amd", "csrc"), "-o", str(exe),
                    str(src)], check=True)
    lines = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    rows = [list(map(int, ln.split())) for ln in lines if ln]
    assert len(rows) == 1003
    for ctr, key, want in KAT:
        assert any(tuple(r[:4]) == ctr and tuple(r[4:6]) == key and tuple(r[6:]) == want for r in rows[:3])
    for r in rows:
        assert tuple(orc.philox(r[:4], r[4:6])) == tuple(r[6:])


def test_photon_shortcuts_on_host(tmp_path):
    """The per-photon shortcuts of the product header (tmc_philox_photon_init
    with tmc_philox4x32_10_block0, _block0_ks, _block0_r2 and _blockb, which take the
    photon's part of rounds 1-2 precomputed) against the full generator
    (pinned above by the vectors), for 20000 pseudo-random (k, b, photon,
    key), b in 0..3: a dropped key, word or round would differ."""
    src = tmp_path / "short.c"
    src.write_text(
        '#include <stdio.h>\n#include "tmc_philox.h"\n'
        "static int same(tmc_u32x4 a, tmc_u32x4 b) { return a.x0 == b.x0 && a.x1 == b.x1 && a.x2 == b.x2 && a.x3 == b.x3; }\n"
        "int main(void){ unsigned s = 7, bad = 0;\n"
        "  for (int i = 0; i < 20000; ++i) { unsigned c[6];\n"
        "    for (int j = 0; j < 6; ++j) { s = s * 1664525u + 1013904223u; c[j] = s; }\n"
        "    const unsigned k = c[0], b = c[1] & 3u, plo = c[2], phi = c[3], k0 = c[4], k1 = c[5];\n"
        "    const tmc_philox_photon pp = tmc_philox_photon_init(plo, phi, k0, k1);\n"
        "    const tmc_philox_keys ks = tmc_philox_key_schedule(k0, k1);\n"
        "    const tmc_u32x4 want0 = tmc_philox4x32_10(k, 0u, plo, phi, k0, k1);\n"
        "    const tmc_u32x4 wantb = tmc_philox4x32_10(k, b, plo, phi, k0, k1);\n"
        "    bad += !same(tmc_philox4x32_10_block0(k, pp, k0, k1), want0);\n"
        "    bad += !same(tmc_philox4x32_10_block0_ks(k, pp, ks), want0);\n"
        "    bad += !same(tmc_philox4x32_10_block0_r2(k, pp, ks), want0);\n"
        "    bad += !same(tmc_philox4x32_10_blockb(k, b, pp, &ks, k0, k1), wantb);\n"
        "  }\n"
        "  printf(\"%u\\n\", bad); return 0; }\n")
    exe = tmp_path / "short"
    subprocess.run(["gcc", "-std=c11", "-O2", "-Wall", "-I", os.path.join(ROOT, "tiny_mc_amd", "csrc"), "-o",
                    str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert out.strip() == "0"


@pytest.mark.gpu
def test_device_philox_kat(tmc, orc):
    rng = np.random.default_rng(7)
    ctr = rng.integers(0, 2**32, size=(4096, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 2**32, size=(4096, 2), dtype=np.uint64).astype(np.uint32)
    for i, (c, k, _) in enumerate(KAT):
        ctr[i] = c
        key[i] = k
    out = tmc.philox(ctr, key)
    for i, (_, _, want) in enumerate(KAT):
        assert tuple(int(v) for v in out[i]) == want
    for i in range(0, 4096, 97):
        assert tuple(int(v) for v in out[i]) == tuple(orc.philox(ctr[i].tolist(), key[i].tolist()))


def test_block0_shortcut_equals_full_rounds(tmp_path):
    """tmc_philox4x32_10_block0 (round 1's photon part precomputed) and its
    form with the round keys precomputed (tmc_philox4x32_10_block0_ks) equal
    the full ten rounds of block (k, 0) on random counters and keys (host
    build of the same header the kernels include)."""
    src = tmp_path / "b0.c"
    src.write_text(r'''
#include <stdio.h>
#include "tmc_philox.h"
static uint64_t s = 88172645463325252ull;
static uint32_t next(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return (uint32_t)(s >> 32); }
int main(void) {
    unsigned bad = 0;
    for (int i = 0; i < 1000000; ++i) {
        uint32_t k = next(), c2 = next(), c3 = next(), k0 = next(), k1 = next();
        if (i & 1) k &= 0xFFFFu;
        tmc_u32x4 a = tmc_philox4x32_10(k, 0u, c2, c3, k0, k1);
        tmc_u32x4 b = tmc_philox4x32_10_block0(k, tmc_philox_photon_init(c2, c3, k0, k1), k0, k1);
        tmc_u32x4 c = tmc_philox4x32_10_block0_ks(k, tmc_philox_photon_init(c2, c3, k0, k1),
                                                  tmc_philox_key_schedule(k0, k1));
        bad += a.x0 != b.x0 || a.x1 != b.x1 || a.x2 != b.x2 || a.x3 != b.x3;
        bad += a.x0 != c.x0 || a.x1 != c.x1 || a.x2 != c.x2 || a.x3 != c.x3;
    }
    printf("%u\n", bad);
    return 0;
}
''')
    exe = tmp_path / "b0"
    inc = os.path.join(ROOT, "tiny_mc_amd", "csrc")
    subprocess.run(["gcc", "-std=c11", "-O2", "-I", inc, "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    assert out.strip() == "0"


def test_walk_block_forms_equal_full_generator(tmp_path):
    """The walk evaluates block (k, b) of a photon from its precomputed part
    of round 1 (tmc_philox4x32_10_block0, _block0_ks and, with any block
    number b, _blockb with stored or formed round keys); each must equal the
    full generator on the same counter and key.  A form that dropped the
    block number, the photon's counter words or a round key would differ."""
    src = tmp_path / "forms.c"
    src.write_text(
        '#include <stdio.h>\n#include "tmc_philox.h"\n'
        "static unsigned long long s = 88172645463325252ull;\n"
        "static unsigned rnd(void){ s ^= s << 13; s ^= s >> 7; s ^= s << 17; return (unsigned)s; }\n"
        "static int same(tmc_u32x4 a, tmc_u32x4 b){ return a.x0==b.x0 && a.x1==b.x1 && a.x2==b.x2 && a.x3==b.x3; }\n"
        "int main(void){ long bad = 0;\n"
        "  for (int i = 0; i < 200000; ++i) {\n"
        "    unsigned k = i % 3 ? rnd() : (unsigned)(i % 1000), b = i % 5 ? (unsigned)(i % 40) : rnd();\n"
        "    unsigned c2 = rnd(), c3 = rnd(), k0 = rnd(), k1 = rnd();\n"
        "    tmc_philox_photon pp = tmc_philox_photon_init(c2, c3, k0, k1);\n"
        "    tmc_philox_keys ks = tmc_philox_key_schedule(k0, k1);\n"
        "    tmc_u32x4 f0 = tmc_philox4x32_10(k, 0, c2, c3, k0, k1), fb = tmc_philox4x32_10(k, b, c2, c3, k0, k1);\n"
        "    bad += !same(f0, tmc_philox4x32_10_block0(k, pp, k0, k1));\n"
        "    bad += !same(f0, tmc_philox4x32_10_block0_ks(k, pp, ks));\n"
        "    bad += !same(fb, tmc_philox4x32_10_blockb(k, b, pp, &ks, k0, k1));\n"
        "    bad += !same(fb, tmc_philox4x32_10_blockb(k, b, pp, 0, k0, k1));\n"
        "    bad += same(fb, tmc_philox4x32_10_blockb(k, b ^ 1u, pp, &ks, k0, k1));\n"
        "  }\n  printf(\"%ld\\n\", bad); return 0; }\n")
    exe = tmp_path / "forms"
    subprocess.run(["gcc", "-std=c11", "-O2", "-I", os.path.join(ROOT, "tiny_mc_amd", "csrc"), "-o", str(exe),
                    str(src)], check=True)
    assert subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.strip() == "0"
