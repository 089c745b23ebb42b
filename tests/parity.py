"""The flight parity gate shared by the GPU parity tests: outcome agreement,
hit-distance tolerance, and the grazing classification of disagreements."""
import numpy as np

from oracle import pyoracle as po

GRAZE = 1e-4  # |f| below which a sign decision is within FP32 evaluation noise (measured <= 2e-6)


EPS_F = 2e-6  # measured bound of |f_gpu - f_oracle| for one evaluation is 1.4e-6


def fragile(F, ray_rec, ref_rec, rel_tol=1e-4, trials=16):
    """True when the oracle's own answer moves beyond the gate under +-EPS_F
    perturbations of its marched values: an ill-conditioned (grazing) flight."""
    for s in range(trials):
        for eps in (EPS_F, -EPS_F):
            r = F.trace_perturbed(ray_rec, eps, 1000 + s)
            if r["outcome"] != ref_rec["outcome"]:
                return True
            if abs(r["t"] - ref_rec["t"]) > rel_tol * abs(ref_rec["t"]) + 1e-6:
                return True
    return False


def parity_gate(got, ref, name, min_agree=0.9999, rel_tol=1e-4, oracle=None):
    """Outcome agreement >= min_agree; every disagreement (outcome, or distance
    beyond rel_tol) must be grazing: the oracle based a sign decision on |f| < GRAZE,
    the secant is ill-conditioned, or `oracle` -- (field, oracle rays) or a
    callable i -> bool -- finds the oracle's own answer fragile."""
    oc = got["outcome"] & 0xFF
    same = oc == ref["outcome"]
    crossing = same & (ref["outcome"] != po.EXIT)
    # hit distance: relative error <= rel_tol, with an absolute floor of 1e-6 (about 4 ulp of a
    # coordinate of magnitude 2) for crossings within a step of the ray origin
    dt = np.abs(got["t"] - ref["t"])
    rel = dt / np.maximum(np.abs(ref["t"]), 1e-30)
    bad = ~same | (crossing & (dt > rel_tol * np.abs(ref["t"]) + 1e-6))
    # Ill-conditioned crossings: the secant t = a + step*pf/(pf - fc) moves by
    # step*(|pf| + |fc|)*eps/(pf - fc)^2 when f moves by eps; with eps = 1e-5 (five
    # times the measured FP32 evaluation error) such a crossing cannot be held to
    # the distance tolerance and counts as grazing too.
    pf, fc = ref["bracket"][:, 0].astype(np.float64), ref["bracket"][:, 1].astype(np.float64)
    step = 0.01
    sens = step * (np.abs(pf) + np.abs(fc)) * 1e-5 / np.maximum((pf - fc) ** 2, 1e-30)
    grazing = (ref["min_abs_f"] < GRAZE) | (sens > rel_tol * np.abs(ref["t"]) + 1e-6)
    if oracle is not None:
        for i in np.nonzero(bad & ~grazing)[0]:
            if callable(oracle):
                grazing[i] = oracle(i)
            else:
                F, orays = oracle
                grazing[i] = fragile(F, orays[i], ref[i], rel_tol)
    n = len(oc)
    print(f"{name}: n={n} agree={1 - bad.mean():.6f} ({bad.sum()} differ, {(bad & ~grazing).sum()} non-grazing); "
          f"crossings={crossing.sum()} max rel t err (agreeing)={rel[crossing & ~bad].max() if (crossing & ~bad).any() else 0:.3e}; "
          f"evals gpu/oracle={got['evals'].sum() / max(ref['evals'].sum(), 1):.4f}")
    for i in np.nonzero(bad)[0][:8]:
        print(f"   ray {i}: gpu ({oc[i]}, t={got['t'][i]:.7f}) oracle ({ref['outcome'][i]}, t={ref['t'][i]:.7f}) "
              f"min|f|={ref['min_abs_f'][i]:.2e} bracket=({pf[i]:.2e},{fc[i]:.2e}) sens={sens[i]:.1e}")
    assert 1 - bad.mean() >= min_agree
    assert not np.any(bad & ~grazing), "non-grazing disagreement"
    ex = (oc == po.EXIT) & (ref["outcome"] == po.EXIT)
    if ex.any():
        assert np.max(np.abs(got["value"][ex] - ref["value"][ex])) < 1e-4
    return bad


