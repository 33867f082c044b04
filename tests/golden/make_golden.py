"""Writes golden.json (and the float enfuse output, enfuse_f32_k3_contrast_entropy.npy)
from the oracle for the cases of cases.py.  Run from the repository root:

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import cases  # noqa: E402
from oracle_binding import enfuse_params, load_oracle  # noqa: E402


def run_enblend(o, c):
    white, wa, black, ba = cases.enblend_inputs(c["seed"], c["itype"], c["w"], c["h"])
    mask, rc = o.nft_mask(wa, ba)
    assert rc == 0
    levels = o.auto_levels(c["w"], c["h"])
    out, oa = o.enblend_pair(white, wa, black, ba, mask, levels, c["wrap"], c["itype"], cases.DITHER_DET, cases.SEED)
    return (white, wa, black, ba), mask, levels, out, oa


def run_enfuse(o, c):
    imgs, alphas = cases.enfuse_inputs(c["seed"], c["itype"], c["K"], c["w"], c["h"], c["holes"])
    levels = o.auto_levels(c["w"], c["h"])
    out, oa = o.enfuse(imgs, alphas, c["itype"], enfuse_params(**c["prm"]), levels, dither=cases.DITHER_DET,
                       seed=cases.SEED)
    return imgs + alphas, levels, out, oa


def run_nft(o, c):
    white, wa, black, ba = cases.enblend_inputs(c["seed"], cases.U8, c["w"], c["h"])
    mask, rc = o.nft_mask(wa, ba)
    dist = o.distance_transform(wa, 2)
    return (wa, ba), mask, rc, dist


def main():
    o = load_oracle()
    rec = {}
    for c in cases.ENBLEND:
        ins, mask, levels, out, oa = run_enblend(o, c)
        rec[c["name"]] = {"inputs": cases.digest(*ins), "levels": levels, "mask": cases.digest(mask),
                          "out": cases.digest(out), "out_alpha": cases.digest(oa)}
    for c in cases.ENFUSE:
        ins, levels, out, oa = run_enfuse(o, c)
        rec[c["name"]] = {"inputs": cases.digest(*ins), "levels": levels, "out": cases.digest(out),
                          "out_alpha": cases.digest(oa)}
        if c["itype"] == cases.F32:
            np.save(os.path.join(HERE, c["name"] + ".npy"), out, allow_pickle=False)
    for c in cases.NFT:
        ins, mask, rc, dist = run_nft(o, c)
        rec[c["name"]] = {"inputs": cases.digest(*ins), "mask": cases.digest(mask), "rc": rc,
                          "distance": cases.digest(dist)}
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(rec, f, indent=1, sort_keys=True)
    print(json.dumps(rec, indent=1))


if __name__ == "__main__":
    main()
