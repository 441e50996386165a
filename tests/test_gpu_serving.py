"""Serving engine on the GPU: continuous batching, admission by HBM budget,
and results that are the same alone and in a mix."""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stream(torch_cuda):
    from relion_amd import serving
    st, cache = serving.make_stream(6, rate_hz=200.0, seed=3, particles=(8, 40),
                                    classes=(1, 3), radius=(2, 4), boxes=(64, 128), order=1,
                                    n_psi=12)
    return st, cache


def test_alone_equals_in_mix(stream, torch_cuda):
    from relion_amd import serving
    st, _ = stream
    probe = st[2][1]
    probe.keep = True
    srv = serving.Server("bf16x3", hbm_budget=8 << 30, pinned_budget=512 << 20, max_inflight=4)
    try:
        tk = srv.submit(probe)
        best_a, idx_a, _ = srv.wait(tk, probe.P)
        alone = srv.diff2(tk, probe.evals())
        srv.release(tk)
        tickets = []
        for k, (_, req) in enumerate(st):
            req.keep = k == 2
            tickets.append(srv.submit(req))
        for k, (_, req) in enumerate(st):
            best, idx, ms = srv.wait(tickets[k], req.P)
            assert ms > 0
            if k == 2:
                mixed = srv.diff2(tickets[k], req.evals())
                srv.release(tickets[k])
                assert np.array_equal(mixed, alone)
                assert np.array_equal(idx, idx_a) and np.array_equal(best, best_a)
        s = srv.stats()
        assert s["completed"] == len(st) + 1 and s["inflight"] == 0 and s["hbm_used"] == 0
        assert s["peak_inflight"] >= 2
    finally:
        srv.close()


def test_server_matches_oracle(stream, torch_cuda):
    from relion_amd import serving, synth
    st, cache = stream
    _, req = st[0]
    req.keep = True
    srv = serving.Server("fp32", hbm_budget=4 << 30, pinned_budget=256 << 20, max_inflight=2)
    try:
        tk = srv.submit(req)
        best, idx, _ = srv.wait(tk, req.P)
        d = srv.diff2(tk, req.evals()).astype(np.float64)
        srv.release(tk)
    finally:
        srv.close()
    y = serving.BOX_TO_COARSE[req.box]
    g = synth.CoarseGeometry(imgY=y, maxR=y // 2)
    off = req.trans_offset
    per = req.C * req.O
    for i in range(min(req.P, 3)):
        tx = req.trans_x[off[i]:off[i + 1]]
        ty = req.trans_y[off[i]:off[i + 1]]
        ref = np.stack([oracle.diff2_f64(cache.volume(req.box, c), g.maxR2_padded,
                                         req.eulers.reshape(-1, 9), tx, ty, req.real[i],
                                         req.imag[i], req.corr[i], req.coor[:, 0],
                                         req.coor[:, 1], n_proj=req.npix)
                        for c in range(req.C)])
        got = d[off[i] * per:off[i + 1] * per]
        st_ = oracle.error_stats(got, ref.ravel())
        assert st_["max_rel"] < 2e-5, st_
        assert idx[i] == oracle.argmin_flat(ref)
        assert abs(best[i] - ref.min()) <= 2e-5 * abs(ref.min())


def test_admission_respects_budget(stream, torch_cuda):
    """With a budget that holds about one large request, requests still all
    complete and the arena never exceeds the budget."""
    from relion_amd import serving
    st, _ = stream
    biggest = max(r.evals() for _, r in st) * 4
    budget = int(biggest * 1.6) + (64 << 20)
    srv = serving.Server("bf16x3", hbm_budget=budget, pinned_budget=512 << 20, max_inflight=8)
    try:
        tickets = [srv.submit(r) for _, r in st]
        for k, (_, r) in enumerate(st):
            srv.wait(tickets[k], r.P)
        s = srv.stats()
        assert s["completed"] == len(st)
        assert s["hbm_peak"] <= budget
    finally:
        srv.close()


def test_request_too_large_is_rejected(stream, torch_cuda):
    from relion_amd import lib, serving
    st, _ = stream
    srv = serving.Server("bf16x3", hbm_budget=2 << 20, pinned_budget=512 << 20, max_inflight=2)
    try:
        with pytest.raises(lib.RcError) as e:
            srv.submit(st[0][1])
        assert e.value.status == 5
    finally:
        srv.close()
