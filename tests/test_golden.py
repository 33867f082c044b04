"""The oracle against the committed golden records of tests/golden (small
seeded enblend, enfuse and NFT cases): inputs regenerate to the recorded
digests and the oracle reproduces the recorded outputs bit for bit."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)

import cases  # noqa: E402
import make_golden  # noqa: E402

with open(os.path.join(HERE, "golden.json")) as _f:
    GOLDEN = json.load(_f)


@pytest.mark.parametrize("c", cases.ENBLEND, ids=lambda c: c["name"])
def test_enblend_golden(oracle, c):
    g = GOLDEN[c["name"]]
    ins, mask, levels, out, oa = make_golden.run_enblend(oracle, c)
    assert cases.digest(*ins) == g["inputs"]
    assert levels == g["levels"]
    assert cases.digest(mask) == g["mask"]
    assert cases.digest(out) == g["out"]
    assert cases.digest(oa) == g["out_alpha"]


@pytest.mark.parametrize("c", cases.ENFUSE, ids=lambda c: c["name"])
def test_enfuse_golden(oracle, c):
    g = GOLDEN[c["name"]]
    ins, levels, out, oa = make_golden.run_enfuse(oracle, c)
    assert cases.digest(*ins) == g["inputs"]
    assert levels == g["levels"]
    assert cases.digest(out) == g["out"]
    assert cases.digest(oa) == g["out_alpha"]
    if c["itype"] == cases.F32:
        np.testing.assert_array_equal(out, np.load(os.path.join(HERE, c["name"] + ".npy"), allow_pickle=False))


@pytest.mark.parametrize("c", cases.NFT, ids=lambda c: c["name"])
def test_nft_golden(oracle, c):
    g = GOLDEN[c["name"]]
    ins, mask, rc, dist = make_golden.run_nft(oracle, c)
    assert cases.digest(*ins) == g["inputs"]
    assert rc == g["rc"]
    assert cases.digest(mask) == g["mask"]
    assert cases.digest(dist) == g["distance"]
