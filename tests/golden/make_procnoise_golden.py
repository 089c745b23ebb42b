"""Writes tests/golden/procnoise.npz from the reference's own simplex noise
(random3, simplex3d, fbm of SdfFunctions.cpp, compiled into oracle/_ref by
oracle/build_ref.sh): the building blocks of the "sandstone" and "rust" ProceduralNoise
fields.  Run in the build container, where the reference tree exists; the committed .npz
travels instead.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import pyoracle as po  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "procnoise.npz")


def main():
    if not po.ref_available():
        raise SystemExit("oracle/_ref is not built (run oracle/build_ref.sh)")
    rng = np.random.default_rng(41)
    # random3 on lattice points (the simplex corners it is called on) and on general points
    r3 = np.concatenate([np.floor(rng.uniform(-300, 300, (1500, 3))), rng.uniform(-50, 50, (500, 3))])
    r3 = r3.astype(np.float32)
    sx = np.concatenate([rng.uniform(-3, 3, (2000, 3)), rng.uniform(-400, 400, (2000, 3))]).astype(np.float32)
    # fbm over the ranges the fields reach: p * 0.3 (sandstone) and p * 2, p * 0.2, p * 50 (rust)
    fb = np.concatenate([rng.uniform(-0.6, 0.6, (1000, 3)), rng.uniform(-4, 4, (1000, 3)),
                         rng.uniform(-100, 100, (500, 3))])
    np.savez_compressed(
        OUT, random3_p=r3, random3=po.random3(r3, "ref"), simplex_p=sx, simplex=po.simplex3d(sx, "ref"),
        fbm_p=fb, fbm2=po.fbm(fb, 2, "ref"), fbm10=po.fbm(fb, 10, "ref"))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
