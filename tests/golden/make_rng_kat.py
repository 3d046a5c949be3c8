"""Writes tests/golden/rng_kat.json: known-answer vectors of the reference RNG
computed with the numpy restatement in rng_numpy.py.

Run: python tests/golden/make_rng_kat.py"""
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import rng_numpy as R  # noqa: E402

STATES = [(0, 0, 0, 0), (1, 2, 3, 4), (255, 255, 0, 0), (1919, 1079, 0xDEADBEEF, 0), (7, 3, 3, 1000),
          (0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFE)]


def main():
    kat = {"pcg4d": [], "rand": [], "tea16": [], "lcg": []}
    for s in STATES:
        kat["pcg4d"].append({"in": list(s), "out": [int(v) for v in R.pcg4d(*s)]})
        vals, w = R.rand_sequence(s, 8)
        kat["rand"].append({"state": list(s), "values_hex": [float(v).hex() for v in vals], "final_w": w})
    for a, b in [(0, 0), (1, 0), (0, 1), (3, 0), (1919, 1079), (0xFFFFFFFF, 12345)]:
        kat["tea16"].append({"in": [a, b], "out": R.tea16(a, b)})
    for seed in [0, 1, 0x9E3779B9, 0xFFFFFFFF]:
        vals, final = R.lcg_sequence(seed, 6)
        kat["lcg"].append({"seed": seed, "values_hex": [float(v).hex() for v in vals], "final": final})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_kat.json")
    with open(path, "w") as f:
        json.dump(kat, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
