"""Writes tests/golden/leaves.npz from the reference's own leaf code
(oracle/_ref/libgpis_ref_leaf.so, compiled by oracle/build_ref.sh from the
reference sources): cell3D/cell1D impulse tables, Box-Muller pairs, PCG output
words and TangentFrame bases.  Run in the build container, where the reference
tree exists; the committed .npz travels to the GPU box instead of the reference.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import pyoracle as po  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "leaves.npz")


def main():
    assert po.ref_available(), "build oracle/_ref first (oracle/build_ref.sh)"
    rng = np.random.default_rng(20251203)
    # 3D cells: a block around the origin for seeds 0, 1, 17429 plus random cells/seeds
    cells, seeds = [], []
    for s in (0, 1, 17429):
        for k in range(-2, 2):
            for j in range(-2, 2):
                for i in range(-2, 2):
                    cells.append((i, j, k)); seeds.append(s)
    for _ in range(256):
        i, j, k = (int(x) for x in rng.integers(-(2 ** 31), 2 ** 31, 3))
        cells.append((i, j, k)); seeds.append(int(rng.integers(0, 2 ** 32)))
    cells = np.array(cells, np.int64)
    seeds = np.array(seeds, np.uint64)
    imp3 = np.stack([po.cell3d(int(c[0]), int(c[1]), int(c[2]), int(s), 8, "ref") for c, s in zip(cells, seeds)])
    imp3_n3 = np.stack([po.cell3d(int(c[0]), int(c[1]), int(c[2]), int(s), 3, "ref") for c, s in zip(cells[:64], seeds[:64])])
    c1 = rng.integers(-(2 ** 31), 2 ** 31, 256).astype(np.int64)
    s1 = rng.integers(0, 2 ** 32, 256).astype(np.uint64)
    imp1 = np.stack([po.cell1d(int(i), int(s), 8, "ref") for i, s in zip(c1, s1)])
    states = rng.integers(0, 2 ** 63, 64).astype(np.uint64)
    bm = np.stack([po.standard_normals(int(s), 4, "ref") for s in states])
    import ctypes as C
    words = np.zeros((64, 32), np.uint32)
    for r, s in enumerate(states):
        po.ref().ref_pcg_words(int(s), 32, words[r].ctypes.data_as(C.POINTER(C.c_uint32)))
    normals = rng.normal(size=(64, 3)).astype(np.float32)
    normals[0] = (0, 0, 1); normals[1] = (0, 0, -1); normals[2] = (1, 0, 0)
    frames = np.stack([po.tangent_frame(n, "ref") for n in normals])
    hx = rng.integers(0, 2 ** 32, (256, 4)).astype(np.uint32)
    h1 = np.array([po.ref().ref_xxhash32_1(int(a)) for a in hx[:, 0]], np.uint32)
    h2 = np.array([po.ref().ref_xxhash32_2(int(a), int(b)) for a, b, _, _ in hx], np.uint32)
    h4 = np.array([po.ref().ref_xxhash32_4(*(int(v) for v in row)) for row in hx], np.uint32)
    np.savez_compressed(OUT, cells3=cells, seeds3=seeds, imp3=imp3, imp3_n3=imp3_n3, cells1=c1, seeds1=s1,
                        imp1=imp1, bm_states=states, bm=bm, pcg_words=words, frame_normals=normals, frames=frames,
                        hash_in=hx, hash1=h1, hash2=h2, hash4=h4)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
