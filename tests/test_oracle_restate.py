"""The oracle's restatement reproduces the unmodified reference byte for byte.

Golden tables were printed by the reference's own sources (tests/golden/
make_golden.py); the restatement (oracle.c, orc_restate_table) walks the same
glibc rand() stream in FP32 in the reference's operation order
(photon.c:6-51) and formats the table as tiny_mc.c:37-66 does.  This pins the
physics text shared with the FP64 oracle (oracle/orc_walk_body.h)."""
import json
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def _manifest():
    with open(os.path.join(GOLDEN, "MANIFEST.json")) as f:
        return json.load(f)


TABLES = sorted(k for k in _manifest()["files"] if k.startswith("ref_"))


@pytest.mark.parametrize("name", TABLES)
def test_restatement_matches_reference_table(orc, name):
    m = _manifest()
    info = m["files"][name]
    phys = m["variants"][info["variant"]]
    p = orc.params(**phys)
    got = orc.restate_table(p, info["seed"], info["photons"])
    with open(os.path.join(GOLDEN, name)) as f:
        want = f.read()
    assert got == want


def test_default_fixture_hash():
    """The seed-12345 table is the unmodified reference's output for
    SEED=12345, PHOTONS=32768, default params.h (tiny_mc.c:37-66 format),
    pinned by its sha256 as taken when the reference was first built."""
    m = _manifest()
    assert (m["files"]["ref_default_s12345_n32768.txt"]["sha256"]
            == "10dcd4cb018ed664be4b74cf859454bda24dc76e7a709c5a4d7ed7da3c382166")


def test_restate_interaction_count(orc):
    """Mean absorptions per photon under the reference's stream: closed form
    73 + 24 p / (1 - p) with p = 0.1 (weight cadence of photon.c:32,45-49)."""
    p = orc.params()
    _, _, n = orc.restate(p, 12345, 32768)
    mean = n / 32768
    assert abs(mean - (73 + 24 * 0.1 / 0.9)) < 0.25
