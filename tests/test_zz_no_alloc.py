"""rc_run_diff2_kernel_coarse_noalloc (scratch in a static device array)
against rc_run_diff2_kernel_coarse (per-stream allocated scratch), through
the C ABI: two calls in a row, the second on a smaller input, each
bit-identical to the allocating entry point on the same data (both run the
same kernels with the same split); sizes beyond the static scratch are
rejected before any launch.  Named to be collected last."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(L):
    return L.rc_run_diff2_kernel_coarse_noalloc


def _t(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a)).cuda()


def test_noalloc_matches_allocating_entry_point(torch_cuda):
    from relion_amd import lib, synth
    torch = torch_cuda
    L = lib.load()
    f = _noalloc(L)
    big = synth.make_problem("c2", particles=1, seed=8)      # 113 x 4608 x 2112 px
    small = synth.make_problem("c1", particles=1, seed=9)    # 21 x 576 x 544 px
    for prob in (big, small):
        pb, g = prob.particles, prob.geom
        order, _ = synth.filter_order(g)
        inv = np.empty_like(order)
        inv[order] = np.arange(order.size)
        proj = lib.Projector(prob.vols[0])
        pk = proj.kernel(g.imgX, g.imgY, g.maxR)
        O = prob.orientation_num
        tx, ty = prob.trans_of(0)
        T = tx.size
        ins = [_t(torch, a) for a in (tx, ty, pb.corr[0][inv], pb.real[0][inv],
                                      pb.imag[0][inv], prob.eulers.ravel())]
        init = np.random.default_rng(1).uniform(0, 3, O * T).astype(np.float32)
        ref = _t(torch, init)
        got = _t(torch, init)
        lib.run_diff2_kernel_coarse(pk, ins[0], ins[1], ins[2], ins[3], ins[4], ins[5], ref, O,
                                    T, g.image_size)
        s = torch.cuda.current_stream().cuda_stream
        st = f(C.byref(pk), ins[0].data_ptr(), ins[1].data_ptr(), None, ins[2].data_ptr(),
               ins[3].data_ptr(), ins[4].data_ptr(), ins[5].data_ptr(), got.data_ptr(), 1.0, O,
               T, g.image_size, s, 0, 0)
        lib.check(st)
        torch.cuda.synchronize()
        print(prob.name, O, T, g.image_size, "split", lib.coarse_split(T, O, g.image_size))
        assert torch.equal(got, ref)
        assert not torch.equal(got, _t(torch, init))
        proj.close()


def test_noalloc_rejects_oversized_shape(torch_cuda):
    """64 translations x 32768 orientations (256 tiles: two pixel slices) over
    a 708 x 1414 half-plane image whose mask (radius 707) holds ~786,000
    pixels: ~17 MB of split partials plus ~15.7 MB of image staging, beyond
    the 24 MiB static scratch: RC_ERR_LIMIT before any launch (the device
    pointers are NULL)."""
    import dataclasses

    from relion_amd import lib, synth
    L = lib.load()
    f = _noalloc(L)
    prob = synth.make_problem("mini", particles=1)
    vol = dataclasses.replace(prob.vols[0], maxR=707)
    proj = lib.Projector(vol)
    imgX, imgY, maxR = 708, 1414, 707
    pk = proj.kernel(imgX, imgY, maxR)
    n = imgX * imgY
    # the drop-in's pixel count: (float)(x^2 + y^2) * pad^2 <= (maxR2_padded + 1)(1 + 2^-12)
    p = np.arange(n)
    x, y = p % imgX, p // imgX
    y = np.where(y > maxR, y - imgY, y)
    pad2 = np.float32(pk.padding_factor) * np.float32(pk.padding_factor)
    thr = (np.float32(pk.maxR2_padded) + np.float32(1)) * np.float32(1 + 2.0 ** -12)
    n_in = int(np.count_nonzero((x * x + y * y).astype(np.float32) * pad2 <= thr))
    need = L.rc_coarse_workspace_bytes(64, 32768, n_in, n_in, 0)
    assert n_in > 700000 and need > 24 << 20, (n_in, need)
    assert lib.coarse_split(64, 32768, n_in)[0] == 2
    st = f(C.byref(pk), None, None, None, None, None, None, None, None, 1.0, 32768, 64, n, None,
           0, 0)
    assert st == 5, st
    assert b"static scratch" in L.rc_last_error()
    proj.close()
