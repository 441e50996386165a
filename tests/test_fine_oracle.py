"""The fine-pass oracle (CPU): cuda_kernel_diff2_fine semantics
(diff2.cuh:540-682) pinned against the coarse float64 oracle where the two
coordinate rules coincide, against a brute-force sum where rows are parked,
and for job-list invariances."""
import numpy as np
import pytest

from oracle import oracle


def make_jobs(sig, max_run):
    """Significant (o, t) mask [O, T] -> job lists: runs of consecutive
    translations of one orientation, at most max_run long."""
    rot, trans, job_idx, job_num = [], [], [], []
    for o in range(sig.shape[0]):
        t = 0
        while t < sig.shape[1]:
            if not sig[o, t]:
                t += 1
                continue
            start = t
            while t < sig.shape[1] and sig[o, t] and t - start < max_run:
                t += 1
            job_idx.append(len(rot))
            job_num.append(t - start)
            for u in range(start, t):
                rot.append(o)
                trans.append(u)
    return (np.asarray(rot, np.int64), np.asarray(trans, np.int64),
            np.asarray(job_idx, np.int64), np.asarray(job_num, np.int64))


@pytest.fixture(scope="module")
def small():
    from relion_amd import synth
    vol = synth.fourier_volume(synth.blob_map(32, seed=2))
    rng = np.random.default_rng(4)
    eul = synth.orientation_set(0, 4)[:24]
    tx, ty, _ = synth.translation_set(5, 32, centre=(0.3, -0.2))
    return vol, rng, eul, tx, ty


def test_fine_equals_coarse_where_rules_agree(small):
    """imgY = 2 maxR: no parked rows, so every (o, t) of the fine pass equals
    the coarse float64 diff2 of the unfiltered image."""
    from relion_amd import synth
    vol, rng, eul, tx, ty = small
    g = synth.CoarseGeometry(imgY=16, maxR=8)
    n = g.imgX * g.imgY
    re, im = rng.normal(size=n).astype(np.float32), rng.normal(size=n).astype(np.float32)
    corr = rng.random(n).astype(np.float32)
    O, T = eul.shape[0], tx.size
    rot, trans, job_idx, job_num = make_jobs(np.ones((O, T), bool), max_run=T)
    fine = oracle.fine_f64(vol, g.maxR2_padded, g.imgX, g.imgY, g.maxR, eul, tx, ty, re, im,
                           corr, rot, trans, job_idx, job_num, sum_init=1.5)
    x, y = synth.pixel_coords(g)
    coarse = oracle.diff2_f64(vol, g.maxR2_padded, eul, tx, ty, re, im, corr, x, y)
    np.testing.assert_allclose(fine.reshape(O, T), coarse + 1.5, rtol=1e-12)


def test_fine_parked_rows_brute_force(small):
    """maxR < imgY/2 - 1: rows between the half-planes sit at x = maxR."""
    from relion_amd import synth
    vol, rng, eul, tx, ty = small
    g = synth.CoarseGeometry(imgY=16, maxR=5)
    n = g.imgX * g.imgY
    re, im = rng.normal(size=n).astype(np.float32), rng.normal(size=n).astype(np.float32)
    corr = rng.random(n).astype(np.float32)
    sig = rng.random((4, tx.size)) < 0.5
    rot, trans, job_idx, job_num = make_jobs(sig, max_run=3)
    got = oracle.fine_f64(vol, g.maxR2_padded, g.imgX, g.imgY, g.maxR, eul[:4], tx, ty, re, im,
                          corr, rot, trans, job_idx, job_num, sum_init=-2.0)
    xs = np.empty(n)
    ys = np.empty(n)
    parked = 0
    for p in range(n):
        x, y = p % g.imgX, p // g.imgX
        if y > g.maxR:
            if y >= g.imgY - g.maxR:
                y -= g.imgY
            else:
                x = g.maxR
                parked += 1
        xs[p], ys[p] = x, y
    assert parked > 0
    for k in range(rot.size):
        P = oracle.project(vol, g.maxR2_padded, eul[rot[k]], xs, ys)
        first = job_idx[np.searchsorted(job_idx, k, side="right") - 1]
        t = trans[first] + (k - first)     # the kernel reads the job's first entry
        assert t == trans[k]
        phi = xs * tx[t] + ys * ty[t]
        X = (re.astype(np.float64) + 1j * im) * np.exp(1j * phi)
        want = -2.0 + np.sum(0.5 * corr * np.abs(P - X) ** 2)
        assert got[k] == pytest.approx(want, rel=1e-12), k


def test_fine_job_list_invariance(small):
    from relion_amd import synth
    vol, rng, eul, tx, ty = small
    g = synth.CoarseGeometry(imgY=16, maxR=6)
    n = g.imgX * g.imgY
    re, im = rng.normal(size=n).astype(np.float32), rng.normal(size=n).astype(np.float32)
    corr = rng.random(n).astype(np.float32)
    sig = rng.random((eul.shape[0], tx.size)) < 0.4
    args = (vol, g.maxR2_padded, g.imgX, g.imgY, g.maxR, eul, tx, ty, re, im, corr)
    a = oracle.fine_f64(*args, *make_jobs(sig, max_run=tx.size))
    b = oracle.fine_f64(*args, *make_jobs(sig, max_run=2))
    rot, trans, job_idx, job_num = make_jobs(sig, max_run=4)
    perm = rng.permutation(job_idx.size)
    c = oracle.fine_f64(*args, rot, trans, job_idx[perm], job_num[perm])
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a, c)
