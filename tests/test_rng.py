"""Oracle RNG (Util/Hash.fs, Base/Sampler.fs) against golden vectors pinned
by the published xxHash32 implementation (tests/golden/gen_rng.py)."""
import json
import os

import numpy as np

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rng.json")))


def test_xxhash32_two(orc):
    for x, y, h in GOLD["xxhash32_two"]:
        assert orc.xxhash2(x, y) == h


def test_xxhash32_three(orc):
    for x, y, z, h in GOLD["xxhash32_three"]:
        assert orc.xxhash3(x, y, z) == h


def test_lcg_stream(orc):
    for rec in GOLD["lcg"]:
        f, end = orc.lcg(rec["seed"], len(rec["floats"]))
        assert np.array_equal(f, np.asarray(rec["floats"], np.float32))
        assert end == rec["end_state"]
        assert np.all((f >= 0) & (f < 1))
