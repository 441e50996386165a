"""Replay of the reference's dump formats: the writer and the C ABI loader
agree (CPU), and a dump pair runs through rc_run_diff2_kernel_coarse to the
oracle's answer (GPU)."""
import numpy as np
import pytest

from oracle import oracle
from relion_amd import lib, replay, synth


def _dump(tmp_path, prob, i):
    pb = prob.particles
    g = prob.geom
    order, _ = synth.filter_order(g)
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size)
    tx, ty = prob.trans_of(i)
    O, T = prob.orientation_num, tx.size
    init = np.random.default_rng(i).uniform(0, 5, O * T).astype(np.float32)
    ppath, dpath = str(tmp_path / "projector_proc_1_1.dat"), str(tmp_path / "coarse_data100_1.dat")
    replay.write_projector_dat(ppath, prob.vols[0])
    replay.write_coarse_dat(dpath, g.imgX, g.imgY, 1, g.maxR, prob.eulers, tx, ty,
                            pb.real[i][inv], pb.imag[i][inv], pb.corr[i][inv], init)
    return ppath, dpath, init, inv


def test_coarse_dat_round_trip(tmp_path):
    prob = synth.make_problem("mini", seed=5)
    ppath, dpath, init, inv = _dump(tmp_path, prob, 1)
    d = replay.load_coarse_dat(dpath)
    g = prob.geom
    assert (d["xdim"], d["ydim"], d["zdim"], d["rmax"]) == (g.imgX, g.imgY, 1, g.maxR)
    assert d["orientation_num"] == prob.orientation_num and d["image_size"] == g.image_size
    np.testing.assert_array_equal(d["eulers"], prob.eulers.ravel())
    np.testing.assert_array_equal(d["real"], prob.particles.real[1][inv])
    np.testing.assert_array_equal(d["diff2s"], init)
    tx, _ = prob.trans_of(1)
    np.testing.assert_array_equal(d["trans_x"], tx)


def test_coarse_dat_truncated(tmp_path):
    prob = synth.make_problem("mini", seed=5)
    _, dpath, _, _ = _dump(tmp_path, prob, 0)
    raw = open(dpath, "rb").read()
    bad = tmp_path / "short.dat"
    bad.write_bytes(raw[:-8])
    with pytest.raises(lib.RcError) as e:
        replay.load_coarse_dat(str(bad))
    assert e.value.status == 6


@pytest.mark.gpu
def test_replay_dump_pair_matches_oracle(tmp_path, torch_cuda):
    prob = synth.make_problem("c1", seed=17, particles=2)
    for i in range(2):
        ppath, dpath, init, inv = _dump(tmp_path, prob, i)
        got, d = replay.replay(ppath, dpath)
        ref = oracle.problem_diff2(prob, i).ravel() + init
        st = oracle.error_stats(got, ref)
        assert st["max_rel"] < 2e-5 and st["mean_rel"] < 1e-6, st
