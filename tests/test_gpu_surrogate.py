"""The learned flight sampler on the GPU against its CPU oracle and against the
plain method it stands in for:

* the fused BF16 MFMA forward/backward kernels give the oracle's gradient (the
  oracle mimics the bf16 rounding points; its own backward is pinned by finite
  differences in test_mlp_oracle.py), the forward gives its class probabilities,
  and one Adam step lands on the oracle's update;
* a short training run on plain-method flights lowers the cross-entropy below
  that of the best predictor without the network's inputs (the marginal class
  frequencies): what the trained network gains over the method without it;
* surrogate flights follow the plain method's distance distribution on fresh rays,
  and every ray that falls back (unconditioned, outside the training hull) gets
  the plain method's reply bit for bit;
* a served flight's endpoint (gradient, outcome by the rejection rules, the value at
  an exit) follows the law of the conditioned process that oracle/served_oracle.py
  restates (pinned on the CPU by tests/test_served_oracle.py).
"""
import math

import numpy as np
import pytest

from oracle import mlp_oracle as M

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _plain_calibration(edges=None):
    from gpis_extend_amd.surrogate import Calibration
    big = np.full(8, 3.0e38, np.float32)
    e = np.linspace(0.0, 1.0, 32).astype(np.float32) if edges is None else edges
    return Calibration(np.zeros(8, np.float32), np.ones(8, np.float32), -big, big, e)


def _net_with(theta):
    from gpis_extend_amd.surrogate import FlightSurrogate
    net = FlightSurrogate(seed=5)
    net.set_params(theta)
    net.set_calibration(_plain_calibration())
    return net


def _layer_slices():
    o = 0
    out = []
    for i, n in M.LAYERS:
        out.append((f"W{len(out) // 2}", slice(o, o + i * n)))
        o += i * n
        out.append((f"b{len(out) // 2}", slice(o, o + n)))
        o += n
    return out


def test_gradient_matches_oracle():
    rng = np.random.default_rng(31)
    theta = M.init_params(seed=3)
    # non-zero biases so every bias path is exercised
    for (W, b) in M.param_views(theta):
        b[...] = rng.normal(0, 0.1, size=b.shape).astype(np.float32)
    net = _net_with(theta)
    B = 256
    x = rng.normal(size=(B, 8)).astype(np.float32)
    y = rng.integers(0, 32, size=B).astype(np.int32)
    g, loss = net.gradient(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda())
    g = g.cpu().numpy().astype(np.float64)
    g_ref = M.gradient(theta, x, y, exact=False)
    l_ref = M.loss(theta, x, y, exact=False)
    print(f"loss gpu {loss:.6f} oracle {l_ref:.6f}")
    assert abs(loss - l_ref) < 2e-3 * max(1.0, l_ref)
    for name, sl in _layer_slices():
        rel = np.linalg.norm(g[sl] - g_ref[sl]) / max(np.linalg.norm(g_ref[sl]), 1e-12)
        print(f"  {name}: |g| {np.linalg.norm(g_ref[sl]):.3e} rel err {rel:.2e}")
        # bf16 operands: a rounding point that lands on the other side of a bf16 tie moves one
        # element by 2^-8 relative; summed over the batch the layer error stays ~1e-3
        assert rel < 2e-2, name
    # and the exact (float64) gradient is within the bf16 quantisation of both
    g_exact = M.gradient(theta, x, y, exact=True)
    assert np.linalg.norm(g - g_exact) / np.linalg.norm(g_exact) < 5e-2
    net.close()


def test_forward_probabilities_match_oracle():
    rng = np.random.default_rng(32)
    theta = M.init_params(seed=7)
    net = _net_with(theta)
    x = rng.normal(size=(1000, 8)).astype(np.float32)  # not a multiple of 64: the tail is masked
    p = net.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    logits, _ = M.forward(theta, x, exact=False)
    p_ref = M.softmax(logits.astype(np.float64))
    err = np.abs(p - p_ref).max()
    print(f"forward max |dp| = {err:.2e}")
    assert err < 5e-3
    assert np.allclose(p.sum(1), 1.0, atol=1e-5)
    net.close()


def test_adam_step_matches_oracle_update():
    rng = np.random.default_rng(33)
    theta = M.init_params(seed=9)
    net = _net_with(theta)
    B = 512
    x = torch.from_numpy(rng.normal(size=(B, 8)).astype(np.float32)).cuda()
    y = torch.from_numpy(rng.integers(0, 32, size=B).astype(np.int32)).cuda()
    g, _ = net.gradient(x, y)
    g = g.cpu().numpy()
    net.train_step(x, y, lr=1e-3)
    got = net.get_params()
    want, _, _ = M.adam_step(theta, np.zeros_like(theta), np.zeros_like(theta), g, 1, lr=1e-3)
    err = np.abs(got - want).max()
    print(f"Adam step max |d theta| = {err:.2e}")
    assert err < 1e-6
    net.close()


def test_parameter_average_is_the_moving_average_of_the_iterates():
    rng = np.random.default_rng(34)
    theta = M.init_params(seed=11)
    net = _net_with(theta)
    beta = 0.9
    net.set_averaging(beta)
    avg = theta.astype(np.float64).copy()
    for k in range(3):
        x = torch.from_numpy(rng.normal(size=(256, 8)).astype(np.float32)).cuda()
        y = torch.from_numpy(rng.integers(0, 32, size=256).astype(np.int32)).cuda()
        net.train_step(x, y, lr=1e-3)
        avg = beta * avg + (1 - beta) * net.get_params().astype(np.float64)
    last = net.get_params()
    net.use_average()
    got = net.get_params()
    print(f"average vs host EMA: max |d| {np.abs(got - avg).max():.2e}; moved from the last iterate "
          f"{np.abs(got - last).max():.2e}")
    assert np.abs(got - avg).max() < 1e-6
    assert np.abs(got - last).max() > 1e-5
    net.close()


@pytest.fixture(scope="module")
def trained():
    """A network trained on 2^22 plain-method flights of the renewal workload (8 epochs,
    512 Adam steps of 65,536), served with the moving average of its iterates (beta 0.99)."""
    from gpis_extend_amd import GpisDevice, SparseConvolutionNoiseMedium, workloads as W
    from gpis_extend_amd import surrogate as S
    med = SparseConvolutionNoiseMedium.from_json(W.config3_json())
    dev = GpisDevice(med)
    net = S.FlightSurrogate(seed=S.SEED_TRAIN)
    feats, s, ex, oc = S.flight_dataset(dev, net, med, 1 << 22, seed=S.SEED_TRAIN)
    cal = S.calibrate(feats, s, ex)
    net.set_calibration(cal)
    cls = S.classes_of(net, s, oc)
    losses = net.train(feats, cls, epochs=8, batch=65536, lr=1e-3, seed=S.SEED_TRAIN, average=0.99)
    counts = torch.bincount(cls[cls >= 0].long(), minlength=32).double()
    pm = counts / counts.sum()
    marginal_ce = float(-(pm[pm > 0] * pm[pm > 0].log()).sum())
    yield {"dev": dev, "net": net, "med": med, "losses": losses, "marginal_ce": marginal_ce, "cal": cal}
    net.close()
    dev.close()


def test_training_gains_over_marginal(trained):
    losses, mce = trained["losses"], trained["marginal_ce"]
    print(f"epoch losses {['%.4f' % v for v in losses]}; marginal (no inputs) {mce:.4f}")
    assert losses[-1] < losses[0]
    # the conditioning state carries information about the next crossing: the trained
    # network must beat the input-free predictor by a clear margin (nats per flight)
    assert losses[-1] < mce - 0.2


def test_surrogate_flights_follow_plain_distribution_per_stratum(trained):
    """The conditional gate: the KS distance between surrogate and plain distances inside
    each of the 27 quantile strata of (value, derivative, length scale), against the same
    statistic between two plain renders with independent realizations (its noise floor).
    An input-free sampler (the pooled law of the plain distances for every ray) passes the
    pooled KS but fails this gate."""
    from gpis_extend_amd import RayBatch, workloads as W
    from gpis_extend_amd import surrogate as S
    dev, net, med = trained["dev"], trained["net"], trained["med"]
    n = 1 << 20
    print(f"case: conditional KS gate, renewal workload, {n} rays seed {S.SEED_INFER}", flush=True)
    rays = W.renewal1d_rays(med, n=n, seed=S.SEED_INFER)
    rb = RayBatch.from_numpy(rays)
    plain = dev.trace_flights_1d(rb)
    rays2 = dict(rays)
    rays2["noise_seed"] = (rays["noise_seed"] ^ np.uint32(0x9E3779B9)).astype(np.uint32)
    rays2["normal_seed"] = (rays["normal_seed"] ^ np.uint32(0x85EBCA6B)).astype(np.uint32)
    rays2["jitter"] = np.random.default_rng(99).uniform(size=n).astype(np.float32)
    rb2 = RayBatch.from_numpy(rays2)
    plain2 = dev.trace_flights_1d(rb2)
    sur, nfb = net.flights(dev, rb)
    fe = net.features(dev, rb)
    torch.cuda.synchronize()
    sp = S.normalised_distance(rb, plain)[0].cpu().numpy()
    sp2 = S.normalised_distance(rb2, plain2)[0].cpu().numpy()
    ss = S.normalised_distance(rb, sur)[0].cpu().numpy()
    sid = S.strata_of(fe.cpu().numpy()[:, list(S.EDGE_STRATA_FEATURES)], 3)
    got = S.stratified_ks(sid, sp, ss)
    floor = S.stratified_ks(sid, sp, sp2)
    free_s = sp[np.random.default_rng(3).permutation(n)]
    free = S.stratified_ks(sid, sp, free_s)
    pooled = S.ks_two_sample(sp, ss)
    print(f"strata KS mean/max: surrogate {got['mean']:.4f}/{got['max']:.4f}, plain vs plain "
          f"{floor['mean']:.4f}/{floor['max']:.4f}, input-free {free['mean']:.4f}/{free['max']:.4f}; "
          f"pooled KS {pooled:.4f}; fallback {nfb}")
    assert got["strata"] == 27
    assert got["mean"] <= floor["mean"] + 0.01 and got["max"] <= floor["max"] + 0.02
    assert S.ks_two_sample(sp, free_s) < 0.01  # the input-free sampler passes the pooled test
    assert free["mean"] > floor["mean"] + 0.05  # ... and fails the conditional one


def test_fallback_rays_get_the_plain_reply(trained):
    from gpis_extend_amd import RayBatch, workloads as W
    dev, net, med = trained["dev"], trained["net"], trained["med"]
    n = 8192
    rays = W.renewal1d_rays(med, n=n, seed=77)
    rng = np.random.default_rng(78)
    # a quarter unconditioned, a quarter with a length scale outside the training hull
    first = rng.uniform(size=n) < 0.25
    rays["flags"] = np.where(first, 1, 0).astype(np.uint32)
    out_hull = (~first) & (rng.uniform(size=n) < 0.33)
    rays["length_scale"] = np.where(out_hull, 0.5, rays["length_scale"]).astype(np.float32)
    rb = RayBatch.from_numpy(rays)
    plain = dev.trace_flights_1d(rb)
    sur, nfb = net.flights(dev, rb)
    torch.cuda.synchronize()
    must = first | out_hull
    print(f"fallback {nfb} of {n} (forced {must.sum()})")
    assert nfb >= must.sum()
    for k in ("t", "value", "gx", "gy", "gz", "outcome", "evals"):
        a = plain[k].cpu().numpy()[must]
        b = sur[k].cpu().numpy()[must]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k
    # surrogate-served rays: a hit or an outward failure inside the bound, or an exit at far
    served = ~must
    oc = sur["outcome"].cpu().numpy() & 0xFF
    t = sur["t"].cpu().numpy()
    ev = sur["evals"].cpu().numpy()
    sv = served & (ev == 0)
    assert sv.sum() > 0.5 * served.sum()
    assert np.all((oc[sv] == 0) | (oc[sv] == 1) | (oc[sv] == 2))
    far = rays["far_t"]
    assert np.all(t[sv & (oc == 0)] == far[sv & (oc == 0)])
    cross = sv & (oc != 0)
    assert np.all((t[cross] >= 0) & (t[cross] <= far[cross]))


def _served_law_inputs(rays, t, lc):
    """Per ray: a, b of the observations, the lag of t, the mean and its slope at t (the
    renewal workload's sphere mean of radius 0.5, sigma 1)."""
    o = np.stack([rays["ox"], rays["oy"], rays["oz"]], 1).astype(np.float64)
    d = np.stack([rays["dx"], rays["dy"], rays["dz"]], 1).astype(np.float64)
    g = np.stack([rays["prev_gx"], rays["prev_gy"], rays["prev_gz"]], 1).astype(np.float64)
    ro = np.linalg.norm(o, axis=1)
    a = rays["prev_value"].astype(np.float64) - (ro - 0.5)
    b = lc * np.einsum("ij,ij->i", g - o / ro[:, None], d)
    q = o + t[:, None].astype(np.float64) * d
    rq = np.linalg.norm(q, axis=1)
    m = rq - 0.5
    ms = np.einsum("ij,ij->i", q / rq[:, None], d)
    return a, b, m, ms, d


def test_served_endpoints_follow_their_law(trained):
    """The endpoint of a served flight: outcomes follow the rejection rules (a hit's
    gradient faces the ray, an outward failure's points along it); a crossing's
    derivative along the ray follows the Rice law of the conditioned process given
    f(t) = 0 on the sense's side (the PIT through oracle/served_oracle.py's CDF is
    uniform); the sense follows the law of the first lattice point's value; the atom at
    the origin returns the observed derivative; an exit's value lies on the sense's
    side with the truncated law; the transverse gradient is the plain method's draw of
    the same seed."""
    from gpis_extend_amd import RayBatch, workloads as W
    from gpis_extend_amd import surrogate as S
    from oracle import served_oracle as so
    from scipy import special
    dev, net, med = trained["dev"], trained["net"], trained["med"]
    n = 1 << 18
    print(f"case: served endpoints, renewal workload, {n} rays seed 61", flush=True)
    rays = W.renewal1d_rays(med, n=n, seed=61)
    rb = RayBatch.from_numpy(rays)
    plain = dev.trace_flights_1d(rb)
    sur, nfb = net.flights(dev, rb)
    torch.cuda.synchronize()
    P = {k: v.cpu().numpy() for k, v in plain.items()}
    R = {k: v.cpu().numpy() for k, v in sur.items()}
    served = R["evals"] == 0
    oc = R["outcome"] & 0xFF
    reason = (R["outcome"] >> 8) & 0xFF
    t = R["t"]
    g = np.stack([R["gx"], R["gy"], R["gz"]], 1).astype(np.float64)
    assert np.all(np.isfinite(g[served])) and np.all(np.isfinite(t[served]))
    lc = rays["length_scale"].astype(np.float64) / np.sqrt(2.0)
    a, b, m, ms, d = _served_law_inputs(rays, t, lc)
    gd = np.einsum("ij,ij->i", g, d)
    hit, fail, ex = served & (oc == 1), served & (oc == 2), served & (oc == 0)
    assert np.all(gd[hit] <= 0) and np.all(gd[fail] > 0) and np.all(reason[fail] == 1)
    # the atom: the observed derivative along the ray
    atom = served & (oc != 0) & (t == 0)
    gobs = (np.stack([rays["prev_gx"], rays["prev_gy"], rays["prev_gz"]], 1).astype(np.float64) * d).sum(1)
    assert np.max(np.abs(gd[atom] - gobs[atom]), initial=0.0) < 1e-4
    # crossings: the Rice law on the sense's side
    cr = served & (oc != 0) & (t > 0)
    e1, e2, c11, c12, c22 = so.cond_law(t[cr] / lc[cr], a[cr], b[cr], True)
    mu, var = so.slope_law(e1, e2, c11, c12, c22, -m[cr])
    muY, sY = ms[cr] + mu / lc[cr], np.sqrt(var) / lc[cr]
    down = oc[cr] == 1
    # a crossing within ~1e-4 correlation lengths of the origin has a slope law of variance
    # 0 in double (the law degenerates to the observed derivative): no PIT there, and such
    # crossings must stay rare
    deg = ~(sY > 0)
    print(f"crossings with a degenerate slope law: {int(deg.sum())}, min t/l {float(np.min(t[cr] / lc[cr])):.2e}")
    assert deg.sum() <= max(20, 1e-4 * len(deg))
    mt = np.where(down, -muY, muY)[~deg] / sY[~deg]
    u = np.sort(so.rice_cdf(np.abs(gd[cr][~deg]) / sY[~deg], mt))
    k = len(u)
    ks = float(np.max(np.maximum(np.arange(1, k + 1) / k - u, u - np.arange(k) / k)))
    print(f"served crossings {k}: Rice PIT KS {ks:.4f}; hits {hit.sum()} fails {fail.sum()} exits {ex.sum()}")
    assert ks < 1.63 / math.sqrt(k) + 0.004
    # the sense: failures among the crossings against the law of f(t0) < 0
    step = np.minimum(np.float32(med.step_size), (rays["far_t"] - rays["near_t"]) / np.float32(med.min_step))
    t0 = rays["near_t"] + step * rays["jitter"]
    a0, b0, m0, _, _ = _served_law_inputs(rays, t0.astype(np.float32), lc)
    e10, _, c110, _, _ = so.cond_law(t0 / lc, a0, b0, True)
    p_neg = special.ndtr((-m0 - e10) / np.sqrt(np.maximum(c110, 1e-300)))
    exp_fail = float(p_neg[cr].sum())
    var_fail = float((p_neg[cr] * (1 - p_neg[cr])).sum())
    print(f"failures among crossings {int(fail[cr].sum())}, expected {exp_fail:.1f}")
    assert abs(fail[cr].sum() - exp_fail) < 5 * math.sqrt(var_fail) + 5
    # exits: the value on the sense's side, distributed as the truncated conditional law
    tf = rays["far_t"]
    af, bf, mf, _, _ = _served_law_inputs(rays, tf, lc)
    e1f, _, c11f, _, _ = so.cond_law(tf / lc, af, bf, True)
    v = R["value"].astype(np.float64)
    z = (v - mf - e1f) / np.sqrt(c11f)
    alpha = (-mf - e1f) / np.sqrt(c11f)
    pos = ex & (v > 0)
    neg = ex & (v < 0)
    assert pos.sum() + neg.sum() == ex.sum()
    up = np.concatenate([so.truncated_normal_cdf_above(z[pos], alpha[pos]),
                         so.truncated_normal_cdf_above(-z[neg], -alpha[neg])])
    up = np.sort(up)
    k = len(up)
    ks = float(np.max(np.maximum(np.arange(1, k + 1) / k - up, up - np.arange(k) / k)))
    print(f"served exits {k}: truncated value law PIT KS {ks:.4f}")
    assert ks < 1.63 / math.sqrt(k) + 0.004
    # the transverse draw is the plain method's: in the ray frame, off the ray direction
    both = served & (P["t"] > 0) & (t > 0) & ((P["outcome"] & 0xFF) != 3)

    def transverse(reply, tt):
        q = np.stack([rays["ox"], rays["oy"], rays["oz"]], 1).astype(np.float64) + tt[:, None] * d
        mg = q / np.linalg.norm(q, axis=1, keepdims=True)
        w = (np.stack([reply["gx"], reply["gy"], reply["gz"]], 1).astype(np.float64) - mg) * lc[:, None]
        return w - np.einsum("ij,ij->i", w, d)[:, None] * d
    idx = np.nonzero(both)[0][:20000]
    ws = transverse(R, t.astype(np.float64))[idx]
    wp = transverse(P, P["t"].astype(np.float64))[idx]
    err = np.abs(ws - wp).max()
    print(f"transverse gradient, served vs plain of the same seed: max |diff| {err:.2e} over {len(idx)} rays")
    assert err < 2e-4
