"""The weights / significance oracle (CPU): RELION's index scan over sorted
cumulative sums (cuda_device_utils.cuh:251-262) against an independent
value-space statement of the same rule, and the weights formula
(utilities.h:405-459) against the identities it must satisfy."""
import numpy as np
import pytest

from oracle import oracle


def value_space_threshold(w, fraction, total):
    """Smallest weight v with sum_{w <= v} w > (1 - fraction) * total (the
    first sorted weight if none), and the count of weights >= v."""
    w = np.asarray(w, np.float64).ravel()
    thr = (1.0 - fraction) * total
    vals = np.unique(w)
    pick = vals[0]
    for v in vals:
        if w[w <= v].sum() > thr:
            pick = v
            break
    return pick, int(np.count_nonzero(w >= pick))


@pytest.mark.parametrize("seed", range(12))
def test_index_scan_equals_value_rule(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 400))
    kind = seed % 4
    if kind == 0:
        w = rng.random(n)
    elif kind == 1:                       # many ties
        w = rng.integers(0, 5, n).astype(np.float64) / 4
    elif kind == 2:                       # exponentially spread, zeros
        w = np.exp(-rng.random(n) * 60)
        w[rng.random(n) < 0.2] = 0.0
    else:                                 # float32-representable weights
        w = rng.random(n).astype(np.float32).astype(np.float64)
    for f in (0.999, 0.9, 0.5, 0.01, 1.0):
        total = float(np.sort(w).cumsum()[-1])
        assert oracle.significance_threshold(w, f, total) == value_space_threshold(w, f, total)


def test_index_scan_edge_cases():
    # all equal: half the weight needs the upper half, every equal weight passes
    ws, n = oracle.significance_threshold(np.ones(10), 0.5)
    assert ws == 1.0 and n == 10
    # fraction 1: the bound is 0, the smallest positive weight is kept
    ws, n = oracle.significance_threshold([0, 0, 0.25, 0.5, 1.0], 1.0)
    assert ws == 0.25 and n == 3
    # a single dominant weight carries the fraction on its own
    ws, n = oracle.significance_threshold([1e-9] * 50 + [1.0], 0.999)
    assert ws == 1.0 and n == 1
    # the bound at the total (no pair): the reference keeps index 0
    ws, n = oracle.significance_threshold([0.5, 1.0], 1.0, total=0.0)
    assert ws == 0.5 and n == 2


def test_weights_formula_identities():
    rng = np.random.default_rng(3)
    C, O, T = 2, 7, 5
    d = rng.random(C * O * T) * 50 + 100
    lpo = rng.normal(size=C * O)
    lpt = rng.normal(size=C * T)
    lpo[3] = -np.inf
    lpt[6] = -np.inf
    mn = d.min()
    w, total, ws, n, lmax = oracle.coarse_weights(d, mn, C, O, T, 0.999, lpo, lpt)
    w3 = w.reshape(C, O, T)
    # excluded samples
    assert np.all(w3[0, 3, :] == 0) and np.all(w3[1, :, 1] == 0)
    # the largest weight is exactly one, the total is the sum
    assert w.max() == 1.0 and total == pytest.approx(w.sum(), rel=1e-15)
    # ratios follow the log weights
    lw = (lpo.reshape(C, O, 1) + lpt.reshape(C, 1, T) + mn - d.reshape(C, O, T))
    a, b = (0, 1, 2), (1, 4, 3)
    assert w3[a] / w3[b] == pytest.approx(np.exp(lw[a] - lw[b]), rel=1e-12)
    # a common shift of diff2 and the minimum changes nothing
    w2, total2, ws2, n2, _ = oracle.coarse_weights(d + 17.0, mn + 17.0, C, O, T, 0.999, lpo, lpt)
    assert np.allclose(w2, w, rtol=1e-12, atol=0) and (ws2, n2) == pytest.approx((ws, n))
    # samples below the (float) minimum are excluded
    d_low = d.copy()
    d_low[0] = mn - 1.0
    w4 = oracle.coarse_weights(d_low, mn, C, O, T, 0.999, lpo, lpt)[0]
    assert w4[0] == 0.0
    # no admissible sample
    wz, tz, wsz, nz, lz = oracle.coarse_weights(d, mn, C, O, T, 0.9, np.full(C * O, -np.inf))
    assert tz == 0.0 and np.all(wz == 0) and lz == -np.inf
