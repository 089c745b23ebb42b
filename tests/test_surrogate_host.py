"""Host logic of the learned flight sampler (no GPU): the minimax bin edges and the
KS helpers used to measure the surrogate."""
import numpy as np

from gpis_extend_amd import surrogate as S


def _linear_gap(samples, edges):
    sc = np.sort(samples)
    xs = np.linspace(0, 1, 20001)
    emp = np.searchsorted(sc, xs, side="right") / len(sc)
    cut = np.searchsorted(sc, edges, side="left") / len(sc)
    return np.max(np.abs(emp - np.interp(xs, edges, cut))[:-1])


def test_minimax_edges_beat_quantile_edges():
    rng = np.random.default_rng(0)
    # a long decaying tail, the shape of normalised crossing distances
    x = np.clip(rng.exponential(0.12, 300000), 0, 1)
    e = S.minimax_edges(x)
    assert len(e) == S.N_BINS + 1 and e[0] == 0.0 and e[-1] == 1.0
    assert np.all(np.diff(e) > 0)
    q = np.quantile(x, np.linspace(0, 1, S.N_BINS + 1))
    q[0], q[-1] = 0.0, 1.0
    gm, gq = _linear_gap(x, e), _linear_gap(x, q)
    assert gm < 0.5 * gq, (gm, gq)
    assert gm < 2e-3


def test_ks_helpers():
    rng = np.random.default_rng(1)
    a = rng.uniform(size=200000)
    b = rng.uniform(size=200000)
    assert S.ks_two_sample(a, b) < 0.01
    assert abs(S.ks_two_sample(a, np.clip(b + 0.05, 0, 1)) - 0.05) < 0.01
    # PIT of uniform samples under a uniform piecewise-linear CDF is uniform
    edges = np.linspace(0, 1, S.N_BINS + 1)
    p = np.full((50000, S.N_CLASS), 0.0)
    p[:, :S.N_BINS] = 1.0 / S.N_BINS
    s = rng.uniform(size=50000)
    assert S.pit_ks(p, edges, s, np.zeros(50000, bool), rng) < 0.015


def test_atom_class_at_origin():
    from oracle import mlp_oracle as M
    edges = np.concatenate([[0.0], np.linspace(0, 1, S.N_BINS)]).astype(np.float32)
    assert edges[0] == edges[1] == 0.0
    s = np.array([0.0, 1e-6, 0.5, 0.99999])
    c = M.class_of(s, np.zeros(4, bool), edges)
    assert c[0] == 0 and c[1] == 1 and c[3] == S.N_BINS - 1
    # sampling the atom returns s = 0 exactly
    p = np.zeros((1000, S.N_CLASS)); p[:, 0] = 0.5; p[:, 5] = 0.5
    rng = np.random.default_rng(2)
    sv, ex = M.cdf_sample(p, edges, rng.uniform(size=1000))
    assert np.all((sv == 0.0) | ((sv >= edges[5]) & (sv <= edges[6])))
    assert 0.4 < np.mean(sv == 0.0) < 0.6


def test_strata_edges_bound_every_stratum_gap():
    """Two strata of different scales (short and long length scales): edges minimax over
    the strata keep each stratum's linearisation gap small, where the pooled edges leave
    the narrow stratum's distribution inside a few wide bins."""
    rng = np.random.default_rng(3)
    narrow = np.clip(rng.gamma(4.0, 0.005, 40000), 0, 1)  # a peaked tenth of the sample
    wide = np.clip(rng.exponential(0.2, 360000), 0, 1)
    x = np.concatenate([narrow, wide])
    lab = np.concatenate([np.zeros(len(narrow), int), np.ones(len(wide), int)])
    e_pool = S.minimax_edges(x, grid=4001)
    e_str = S.minimax_edges(x, grid=4001, strata=lab)
    assert len(e_str) == S.N_BINS + 1 and np.all(np.diff(e_str) > 0)
    g_pool = max(_linear_gap(narrow, e_pool), _linear_gap(wide, e_pool))
    g_str = max(_linear_gap(narrow, e_str), _linear_gap(wide, e_str))
    assert g_str < 0.6 * g_pool, (g_str, g_pool)
    assert g_str < 5e-3


def test_stratified_ks_and_image_bias():
    rng = np.random.default_rng(4)
    n = 200000
    key = rng.uniform(size=(n, 1))
    a = rng.exponential(1.0, n) * (0.5 + key[:, 0])
    b = rng.exponential(1.0, n) * (0.5 + key[:, 0])
    sid = S.strata_of(key, 4)
    same = S.stratified_ks(sid, a, b)
    # an input-free sampler (the pooled law for every row) fails the conditional test
    free = S.stratified_ks(sid, a, b[rng.permutation(n)])
    assert same["strata"] == 4 and same["mean"] < 0.01
    assert free["mean"] > 5 * same["mean"] and free["mean"] > 0.05
    # image bias: zero in expectation for two unbiased renders, the bias when one is shifted
    px, spp = 4096, 16
    mu = rng.uniform(0.2, 1.0, (px, 1))
    A = mu + rng.normal(0, 0.3, (px, spp))
    B = mu + rng.normal(0, 0.3, (px, spp))
    rms, limit = S.image_bias(A, B)
    assert rms <= limit < 0.05, (rms, limit)
    rms, limit = S.image_bias(A * 1.1, B)
    want = 0.1 * np.sqrt(np.mean(mu ** 2)) / np.mean(mu)
    assert rms > limit and abs(rms - want) < 0.2 * want, (rms, want, limit)
