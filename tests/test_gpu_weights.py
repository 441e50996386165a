"""rc_coarse_weights (GPU) against the float64 oracle: weights, their sum,
the significance threshold and mask (utilities.h:405-459,
cuda_device_utils.cuh:238-267 semantics)."""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


def _search(lib, torch, prob, form="fp32"):
    projs = [lib.Projector(v) for v in prob.vols]
    plan = lib.plan_for_problem(prob, projs, form)
    t = lib.problem_tensors(prob)
    plan.execute(t)
    torch.cuda.synchronize()
    return t


def _priors(prob, rng, excluded=0.05):
    C, O = prob.n_classes, prob.orientation_num
    lpo = rng.normal(scale=2.0, size=C * O).astype(np.float32)
    lpo[rng.random(lpo.size) < excluded] = -np.inf
    lpt = rng.normal(scale=1.0, size=C * int(prob.trans_offset[-1])).astype(np.float32)
    lpt[rng.random(lpt.size) < excluded] = -np.inf
    return lpo, lpt


def _check(prob, t, out, fraction, lpo, lpt):
    _check_arrays(t["diff2"].cpu().numpy(), t["best_diff2"].cpu().numpy(),
                  np.asarray(prob.trans_offset, np.int64), prob.n_classes,
                  prob.orientation_num, out, fraction, lpo, lpt)


def _check_arrays(d, best, off, C, O, out, fraction, lpo, lpt):
    w = out["weights"].cpu().numpy()
    mask = out["mask"].cpu().numpy()
    S = out["sum"].cpu().numpy()
    ws = out["significant_weight"].cpu().numpy()
    ns = out["n_significant"].cpu().numpy()
    for i in range(len(off) - 1):
        T = int(off[i + 1] - off[i])
        a, b = C * O * off[i], C * O * off[i + 1]
        pt = None if lpt is None else lpt[C * off[i]:C * off[i + 1]]
        rw, rS, rws, rn, _ = oracle.coarse_weights(d[a:b], best[i], C, O, T, fraction, lpo, pt)
        gw = w[a:b].astype(np.float64)
        big = rw > 1e-30
        assert np.all(np.abs(gw[big] - rw[big]) <= 2e-7 * rw[big]), i
        assert np.all(np.abs(gw[~big]) <= 1e-30), i
        assert S[i] == pytest.approx(rS, rel=2e-7)   # float-stored weights
        assert S[i] == pytest.approx(gw.sum(), rel=1e-12)
        # the selection rule applied to the GPU's own weights gives its answer
        gws, gn = oracle.significance_threshold(gw, fraction, S[i])
        assert (np.float32(gws), gn) == (ws[i], ns[i]), (i, gws, gn, ws[i], ns[i])
        assert np.array_equal(mask[a:b], (w[a:b] >= ws[i]).astype(np.uint8))
        # and the oracle's float64 weights select (almost) the same samples
        assert abs(int(ns[i]) - rn) <= max(2, rn // 1000), (i, ns[i], rn)


@pytest.fixture(scope="module")
def lib(torch_cuda):
    from relion_amd import lib as L
    L.load()
    return L


@pytest.mark.parametrize("fraction", [0.999, 0.9, 1.0])
def test_weights_c1(lib, torch_cuda, fraction):
    from relion_amd import synth
    torch = torch_cuda
    prob = synth.make_problem("c1", seed=11)
    t = _search(lib, torch, prob)
    rng = np.random.default_rng(5)
    lpo, lpt = _priors(prob, rng)
    dlpo = torch.as_tensor(lpo).cuda()
    dlpt = torch.as_tensor(lpt).cuda()
    out = lib.coarse_weights(prob.trans_offset, t["trans_offset"], 1, prob.orientation_num,
                             t["diff2"], t["best_diff2"], dlpo, 0, dlpt, fraction)
    torch.cuda.synchronize()
    _check(prob, t, out, fraction, lpo, lpt)


def test_weights_ragged_multiclass_no_priors_in_place(lib, torch_cuda):
    """Two classes, uneven translation counts, no priors, weights written
    over diff2; a second run and a one-particle batch agree bit for bit."""
    from relion_amd import synth
    torch = torch_cuda
    vols = [synth.fourier_volume(synth.blob_map(32, seed=s)) for s in (3, 4)]
    g = synth.CoarseGeometry(imgY=16, maxR=8)
    eul = synth.orientation_set(0, 8)
    pb = synth.make_particles(vols, g, 5, 32, 0.5, seed=9, shift_range=1.0)
    txs, tys, offs = [], [], [0]
    for i, r2 in enumerate((1, 13, 4, 2, 0)):
        tx, ty, _ = synth.translation_set(r2, 32, centre=(0.5 * i, -0.25 * i))
        txs.append(tx)
        tys.append(ty)
        offs.append(offs[-1] + tx.size)
    prob = synth.CoarseProblem("ragged", vols, g, eul, np.asarray(offs, np.int32),
                               np.concatenate(txs), np.concatenate(tys), pb, 32)
    t = _search(lib, torch, prob)
    keep = t["diff2"].clone()
    out = lib.coarse_weights(prob.trans_offset, t["trans_offset"], 2, prob.orientation_num,
                             keep.clone(), t["best_diff2"], fraction=0.99)
    inplace = keep.clone()
    out2 = lib.coarse_weights(prob.trans_offset, t["trans_offset"], 2, prob.orientation_num,
                              inplace, t["best_diff2"], fraction=0.99, weights=inplace)
    torch.cuda.synchronize()
    _check(prob, t, out, 0.99, None, None)
    for k in ("weights", "sum", "significant_weight", "n_significant", "mask"):
        assert torch.equal(out[k], out2[k]), k
    # particle 3 alone
    i = 3
    a, b = 2 * prob.orientation_num * offs[i], 2 * prob.orientation_num * offs[i + 1]
    one_off = np.array([0, offs[i + 1] - offs[i]], np.int32)
    one = lib.coarse_weights(one_off, torch.as_tensor(one_off).cuda(), 2, prob.orientation_num,
                             keep[a:b].clone(), t["best_diff2"][i:i + 1].clone(), fraction=0.99)
    torch.cuda.synchronize()
    assert torch.equal(one["weights"], out["weights"][a:b])
    assert torch.equal(one["significant_weight"], out["significant_weight"][i:i + 1])
    assert torch.equal(one["n_significant"], out["n_significant"][i:i + 1])


def test_weights_rejects_bad_fraction(lib, torch_cuda):
    torch = torch_cuda
    off = np.array([0, 2], np.int32)
    x = torch.zeros(4, device="cuda")
    m = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(lib.RcError):
        lib.coarse_weights(off, torch.as_tensor(off).cuda(), 1, 2, x, m, fraction=0.0)


def test_weights_crowded_buckets_and_alignment(lib, torch_cuda):
    """Particles whose selected first-level bucket (exponent and three leading
    significand bits) holds more samples than the kernel finishes in LDS --
    near-equal and equal weights -- beside ones that fit; sizes that put the
    particle boundaries inside 16-byte lines; weights and mask buffers at
    another alignment than diff2.  The outputs match those of aligned buffers
    bit for bit, and a run without a mask gives the same threshold and count."""
    torch = torch_cuda
    rng = np.random.default_rng(21)
    C, O = 1, 96
    Ts = [70, 53, 80, 1, 45, 3]
    off = np.concatenate([[0], np.cumsum(Ts)]).astype(np.int64)
    parts, best = [], []
    for i, T in enumerate(Ts):
        n = C * O * T
        base = 1000.0 + 10 * i
        if i == 0:
            x = base + rng.random(n) * 0.01            # weights in [0.99, 1]
        elif i == 1:
            x = np.full(n, base)                        # every weight 1
        elif i == 2:                                    # 70 % crowded, rest spread
            x = base + np.where(rng.random(n) < 0.7, rng.random(n) * 0.02, rng.random(n) * 50)
        elif i == 3:
            x = base + rng.random(n) * 3
        else:
            x = base + rng.random(n) * 200
        x = x.astype(np.float32)
        parts.append(x)
        best.append(float(x.min()))
    d = np.concatenate(parts)
    best = np.asarray(best)
    d_t = torch.as_tensor(d).cuda()
    best_t = torch.as_tensor(best).cuda()
    off32 = off.astype(np.int32)
    doff = torch.as_tensor(off32).cuda()
    n = d.size
    for fraction in (0.999, 0.5):
        ref = lib.coarse_weights(off32, doff, C, O, d_t, best_t, fraction=fraction)
        wbuf = torch.full((n + 1,), -1.0, device="cuda")
        mbuf = torch.full((n + 3,), 7, dtype=torch.uint8, device="cuda")
        odd = lib.coarse_weights(off32, doff, C, O, d_t, best_t, fraction=fraction,
                                 weights=wbuf[1:], mask=mbuf[3:])
        nomask = lib.coarse_weights(off32, doff, C, O, d_t, best_t, fraction=fraction,
                                    mask=False)
        torch.cuda.synchronize()
        _check_arrays(d, best, off, C, O, ref, fraction, None, None)
        for k in ("weights", "sum", "significant_weight", "n_significant", "mask"):
            assert torch.equal(ref[k], odd[k]), (fraction, k)
        for k in ("weights", "sum", "significant_weight", "n_significant"):
            assert torch.equal(ref[k], nomask[k]), (fraction, k)
        assert wbuf[0].item() == -1.0 and mbuf[:3].tolist() == [7, 7, 7]


def test_weights_minimum_not_attained(lib, torch_cuda):
    """Without priors the kernel takes min - stored(min) as the largest log
    weight and repeats its weights pass when no sample attains it: a minimum
    below every sample (largest log weight smaller than the bound) and one
    above some samples (those get weight 0), each against the oracle, next to
    the minimum the search reports."""
    from relion_amd import synth
    torch = torch_cuda
    prob = synth.make_problem("c1", seed=13, particles=4)
    t = _search(lib, torch, prob)
    best = t["best_diff2"].clone()
    d = t["diff2"].cpu().numpy()
    off = np.asarray(prob.trans_offset, np.int64)
    O = prob.orientation_num
    for shift in (0.0, -0.37, 2.5):
        m = best + shift
        out = lib.coarse_weights(prob.trans_offset, t["trans_offset"], 1, O, t["diff2"], m,
                                 fraction=0.999)
        torch.cuda.synchronize()
        _check_arrays(d, m.cpu().numpy(), off, 1, O, out, 0.999, None, None)
        if shift > 0:   # samples under the minimum are excluded
            w = out["weights"].cpu().numpy()
            for i in range(prob.n_particles):
                a, b = O * off[i], O * off[i + 1]
                below = d[a:b] < np.float32(m[i].item())
                assert below.any() and np.all(w[a:b][below] == 0.0)
