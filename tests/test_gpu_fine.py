"""rc_run_diff2_kernel_fine (GPU) against the float64 fine-pass oracle
(cuda_kernel_diff2_fine, diff2.cuh:540-682 semantics)."""
import numpy as np
import pytest

from oracle import oracle
from test_fine_oracle import make_jobs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(torch_cuda):
    from relion_amd import lib as L
    L.load()
    return L


def _run(lib, torch, vol, g, eul, tx, ty, re, im, corr, jobs, sum_init, n_out, fill=np.nan):
    rot, trans, job_idx, job_num = jobs
    t = lambda a, dt=torch.float32: torch.as_tensor(np.ascontiguousarray(a)).to("cuda", dt)
    proj = lib.Projector(vol)
    pk = proj.kernel(g.imgX, g.imgY, g.maxR)
    out = torch.full((n_out,), float(fill), dtype=torch.float32, device="cuda")
    lib.run_diff2_kernel_fine(pk, t(eul.ravel()), t(re), t(im), t(tx), t(ty), t(corr), out,
                              re.size, sum_init, eul.shape[0], tx.size, t(rot, torch.int64),
                              t(trans, torch.int64), t(job_idx, torch.int64),
                              t(job_num, torch.int64))
    torch.cuda.synchronize()
    res = out.cpu().numpy()
    proj.close()
    return res


def test_fine_c1_random_jobs(lib, torch_cuda):
    """Particle images of the 32^2 workload, 30 % of (o, t) significant, runs
    of up to 20 translations (more than one 16-translation group)."""
    from relion_amd import synth
    torch = torch_cuda
    prob = synth.make_problem("c1", seed=11)
    g = prob.geom
    order, _ = synth.filter_order(g)
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size)
    pb = prob.particles
    rng = np.random.default_rng(8)
    for i in (0, 5):
        tx, ty = prob.trans_of(i)
        sig = rng.random((prob.orientation_num, tx.size)) < 0.3
        sig[7, :] = True                       # one full row: a 21-translation run
        jobs = make_jobs(sig, max_run=20)
        re, im, corr = pb.real[i][inv], pb.imag[i][inv], pb.corr[i][inv]
        got = _run(lib, torch, prob.vols[0], g, prob.eulers, tx, ty, re, im, corr, jobs, 3.0,
                   jobs[0].size)
        ref = oracle.fine_f64(prob.vols[0], g.maxR2_padded, g.imgX, g.imgY, g.maxR, prob.eulers,
                              tx, ty, re, im, corr, *jobs, sum_init=3.0)
        st = oracle.error_stats(got, ref)
        print(f"fine c1 particle {i}: {jobs[3].size} jobs, {st}")
        assert st["max_rel"] <= 2e-5 and st["mean_rel"] <= 1e-6, st


def test_fine_parked_rows_and_skipped_job(lib, torch_cuda):
    from relion_amd import synth
    torch = torch_cuda
    vol = synth.fourier_volume(synth.blob_map(32, seed=2))
    rng = np.random.default_rng(4)
    eul = synth.orientation_set(0, 4)[:24]
    tx, ty, _ = synth.translation_set(5, 32, centre=(0.3, -0.2))
    g = synth.CoarseGeometry(imgY=16, maxR=5)
    n = g.imgX * g.imgY
    re, im = rng.normal(size=n).astype(np.float32), rng.normal(size=n).astype(np.float32)
    corr = rng.random(n).astype(np.float32)
    sig = rng.random((eul.shape[0], tx.size)) < 0.5
    rot, trans, job_idx, job_num = make_jobs(sig, max_run=6)
    # a job naming an orientation past the array is skipped
    bad = job_idx.size - 1
    rot_bad = rot.copy()
    rot_bad[job_idx[bad]] = eul.shape[0] + 5
    got = _run(lib, torch, vol, g, eul, tx, ty, re, im, corr,
               (rot_bad, trans, job_idx, job_num), 0.0, rot.size, fill=-7.0)
    ref = oracle.fine_f64(vol, g.maxR2_padded, g.imgX, g.imgY, g.maxR, eul, tx, ty, re, im, corr,
                          rot, trans, job_idx, job_num)
    sel = np.ones(rot.size, bool)
    sel[job_idx[bad]:job_idx[bad] + job_num[bad]] = False
    st = oracle.error_stats(got[sel], ref[sel])
    assert st["max_rel"] <= 2e-5 and st["mean_rel"] <= 1e-6, st
    assert np.all(got[~sel] == -7.0)


def test_fine_rejects_3d_and_accepts_no_jobs(lib, torch_cuda):
    import ctypes as C
    torch = torch_cuda
    from relion_amd import synth
    vol = synth.fourier_volume(synth.blob_map(16, seed=1))
    proj = lib.Projector(vol)
    pk = proj.kernel(9, 16, 8)
    L = lib.load()
    st = L.rc_run_diff2_kernel_fine(C.byref(pk), *([None] * 8), 100, 0.0, 1, 1, 1,
                                    *([None] * 4), 1, None)
    assert st == 2
    z = torch.zeros(1, device="cuda")
    st = L.rc_run_diff2_kernel_fine(C.byref(pk), *([z.data_ptr()] * 8), 100, 0.0, 1, 1, 0,
                                    *([z.data_ptr()] * 4), 0, None)
    assert st == 0
    proj.close()
