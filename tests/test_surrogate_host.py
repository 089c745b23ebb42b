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
