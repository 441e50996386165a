"""Single-image entry points on the GPU at a production particle shape
(113 translations x 4608 orientations x 1636 in-radius pixels): the split-K
work division against the float64 oracle, repeatability, per-stream scratch
under concurrent streams, the device in-radius partition, and the
autotuner's choice reaching later plans."""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

# max / mean relative error against float64 per form: about 5x the error
# measured at this shape (FP32 1.7e-7 / 3.3e-8, BF16x3 3.3e-7 / 4.8e-8,
# FP16 1.5e-5 / 2.3e-6)
TOL = {"fp32": (1e-6, 2e-7), "bf16x3": (1.7e-6, 2.5e-7), "fp16": (7.5e-5, 1.2e-5)}


@pytest.fixture(scope="module")
def c2p():
    from relion_amd import synth
    prob = synth.make_problem("c2", particles=4, seed=5)
    ref = oracle.problem_diff2_expanded(prob, range(4))
    return prob, ref


@pytest.fixture(scope="module")
def lib(torch_cuda):
    from relion_amd import lib as L
    L.load()
    return L


def _t(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a)).cuda()


def _inverse(order):
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size)
    return inv


def _check(got, ref, form, what):
    st = oracle.error_stats(got, ref)
    mx, mean = TOL[form]
    print(f"{what} [{form}] max_rel={st['max_rel']:.3e} mean_rel={st['mean_rel']:.3e}")
    assert st["max_rel"] <= mx and st["mean_rel"] <= mean, (what, form, st)


@pytest.mark.parametrize("form", ["fp32", "bf16x3", "fp16"])
def test_split_single_production_particle(c2p, lib, torch_cuda, form):
    """rc_coarse_run on one 113 x 4608 x 1636 particle: the pixel axis is
    split (several slices), the result matches float64, and two calls give
    the same bits."""
    torch = torch_cuda
    prob, ref = c2p
    pb, g = prob.particles, prob.geom
    O = prob.orientation_num
    ks, kc = lib.coarse_split(113, O, pb.npix, form)
    assert ks > 1 and (ks - 1) * kc < (pb.npix + 31) // 32 <= ks * kc, (ks, kc)
    proj = lib.Projector(prob.vols[0])
    pk = proj.kernel(g.imgX, g.imgY, g.maxR)
    eul, coor = _t(torch, prob.eulers.ravel()), _t(torch, pb.coor_xy.ravel())
    k = lib.CoarseKernel(113, O, pb.npix, pb.full_npix, form)
    ws = torch.empty(k.get_workspace_size_bytes(), dtype=torch.uint8, device="cuda")
    for i in (0, 3):
        tx, ty = prob.trans_of(i)
        args = (eul, _t(torch, tx), _t(torch, ty), _t(torch, pb.real[i]), _t(torch, pb.imag[i]),
                pk, _t(torch, pb.corr[i]))
        outs = []
        for _ in range(2):
            dst = torch.full((O * 113,), float("nan"), dtype=torch.float32, device="cuda")
            k.run(*args, None, dst, coor_xy=coor, workspace=ws)
            outs.append(dst)
        torch.cuda.synchronize()
        assert torch.equal(outs[0], outs[1])
        got = outs[0].cpu().numpy().astype(np.float64)
        _check(got, ref[i].ravel(), form, f"split particle {i} ({ks} slices)")
        if form != "fp16":
            assert int(np.argmin(got)) == oracle.argmin_flat(ref[i])


def test_drop_in_production_particle_against_direct_oracle(c2p, lib, torch_cuda):
    """rc_run_diff2_kernel_coarse (unfiltered image, in-place +=) on one
    production particle against the direct-form float64 oracle."""
    from relion_amd import synth
    torch = torch_cuda
    prob, _ = c2p
    pb, g = prob.particles, prob.geom
    inv = _inverse(synth.filter_order(g)[0])
    O = prob.orientation_num
    i = 1
    tx, ty = prob.trans_of(i)
    ref = oracle.problem_diff2(prob, i).ravel()
    proj = lib.Projector(prob.vols[0])
    pk = proj.kernel(g.imgX, g.imgY, g.maxR)
    d2 = torch.full((O * 113,), 2.5, dtype=torch.float32, device="cuda")
    lib.run_diff2_kernel_coarse(pk, _t(torch, tx), _t(torch, ty), _t(torch, pb.corr[i][inv]),
                                _t(torch, pb.real[i][inv]), _t(torch, pb.imag[i][inv]),
                                _t(torch, prob.eulers.ravel()), d2, O, 113, g.image_size)
    torch.cuda.synchronize()
    _check(d2.cpu().numpy(), ref + 2.5, "fp32", "drop-in particle 1")


def test_drop_in_on_concurrent_streams(c2p, lib, torch_cuda):
    """Two images on two streams, launched back to back without waiting:
    each stream has its own scratch, so each result equals the same image
    run alone."""
    from relion_amd import synth
    torch = torch_cuda
    prob, _ = c2p
    pb, g = prob.particles, prob.geom
    inv = _inverse(synth.filter_order(g)[0])
    O = prob.orientation_num
    proj = lib.Projector(prob.vols[0])
    pk = proj.kernel(g.imgX, g.imgY, g.maxR)
    eul = _t(torch, prob.eulers.ravel())
    ins = []
    for i in (0, 2):
        tx, ty = prob.trans_of(i)
        ins.append((_t(torch, tx), _t(torch, ty), _t(torch, pb.corr[i][inv]),
                    _t(torch, pb.real[i][inv]), _t(torch, pb.imag[i][inv])))
    solo = []
    for tx, ty, c, re, im in ins:
        d = torch.zeros(O * 113, dtype=torch.float32, device="cuda")
        lib.run_diff2_kernel_coarse(pk, tx, ty, c, re, im, eul, d, O, 113, g.image_size)
        solo.append(d)
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for rep in range(3):
        outs = [torch.zeros(O * 113, dtype=torch.float32, device="cuda") for _ in ins]
        torch.cuda.synchronize()
        for j in (0, 1):
            tx, ty, c, re, im = ins[j]
            lib.run_diff2_kernel_coarse(pk, tx, ty, c, re, im, eul, outs[j], O, 113,
                                        g.image_size, stream=streams[j])
        torch.cuda.synchronize()
        for j in (0, 1):
            assert torch.equal(outs[j], solo[j]), (rep, j)
    for s in streams:
        lib.scratch_release(s)


def test_drop_in_projection_cache(c2p, lib, torch_cuda):
    """The drop-in keeps, per stream, the projections of its last call: calls
    with the same projector and orientations reuse them and give the bits of
    the allocation-free entry point, which gathers on every call; new
    orientations written into the same buffer, and another projector, are
    caught and give that entry point's bits for them."""
    from relion_amd import synth
    torch = torch_cuda
    prob, _ = c2p
    pb, g = prob.particles, prob.geom
    inv = _inverse(synth.filter_order(g)[0])
    O, i = prob.orientation_num, 3
    tx, ty = prob.trans_of(i)
    ins = [_t(torch, a) for a in (tx, ty, pb.corr[i][inv], pb.real[i][inv], pb.imag[i][inv])]
    import dataclasses
    v0 = prob.vols[0]
    half = dataclasses.replace(v0, real=v0.real * np.float32(0.5), imag=v0.imag * np.float32(0.5))
    projs = [lib.Projector(v0), lib.Projector(half)]
    pks = [p.kernel(g.imgX, g.imgY, g.maxR) for p in projs]
    eul = _t(torch, prob.eulers.ravel())
    s = torch.cuda.Stream()

    def run(pk, fn, stream=None):
        d = torch.zeros(O * 113, dtype=torch.float32, device="cuda")
        fn(pk, *ins, eul, d, O, 113, g.image_size, stream=stream)
        torch.cuda.synchronize()
        return d

    cached = lambda pk: run(pk, lib.run_diff2_kernel_coarse, s)
    gathered = lambda pk: run(pk, lib.run_diff2_kernel_coarse_noalloc)
    ref = gathered(pks[0])
    for k in range(3):
        assert torch.equal(cached(pks[0]), ref), k
    # other orientations in the same buffer: the eulers rotated by one row
    eul.copy_(_t(torch, np.roll(prob.eulers, 1, axis=0).ravel()))
    ref2 = gathered(pks[0])
    assert not torch.equal(ref2, ref)
    assert torch.equal(cached(pks[0]), ref2)
    assert torch.equal(cached(pks[0]), ref2)
    # another projector, then the first one again
    ref3 = gathered(pks[1])
    assert torch.equal(cached(pks[1]), ref3)
    assert torch.equal(cached(pks[0]), ref2)
    lib.scratch_release(s)


def test_filter_images_on_device(lib, torch_cuda):
    """rc_filter_coarse_images reproduces the host order exactly: the test
    problems' filtered images from their unfiltered form, and random data
    under a non-integer padding factor."""
    from relion_amd import synth
    torch = torch_cuda
    prob = synth.make_problem("c1", seed=11)
    pb, g = prob.particles, prob.geom
    n = g.image_size
    inv = _inverse(synth.filter_order(g)[0])
    fimg = np.concatenate([pb.real[:, inv], pb.imag[:, inv]], axis=1)      # [P][2n]
    out = lib.filter_images(lib.geometry_view(g.imgX, g.imgY, g.maxR, g.pad), n,
                            _t(torch, fimg), _t(torch, pb.corr[:, inv]))
    torch.cuda.synchronize()
    assert np.array_equal(out["real"].cpu().numpy(), pb.real)
    assert np.array_equal(out["imag"].cpu().numpy(), pb.imag)
    assert np.array_equal(out["corr"].cpu().numpy(), pb.corr)
    assert np.array_equal(out["coor_xy"].cpu().numpy(), pb.coor_xy.ravel())
    assert np.all(out["filtered_size"].cpu().numpy() == pb.npix)
    # 40-row image, maxR 17, pad 1.7, three images of random values, 3000 px
    rng = np.random.default_rng(4)
    view = lib.geometry_view(21, 40, 17, 1.7, int(17 * 17 * np.float32(1.7) ** 2))
    n = 800
    order, n_in = lib.filter_order(view, n)
    f = rng.standard_normal((3, 2 * n)).astype(np.float32)
    c = rng.random((3, n)).astype(np.float32)
    out = lib.filter_images(view, n, _t(torch, f), _t(torch, c))
    torch.cuda.synchronize()
    assert np.array_equal(out["real"].cpu().numpy(), f[:, :n][:, order])
    assert np.array_equal(out["imag"].cpu().numpy(), f[:, n:][:, order])
    assert np.array_equal(out["corr"].cpu().numpy(), c[:, order])
    i = order
    x, y = i % 21, i // 21
    y = np.where(y > 17, y - 40, y)
    assert np.array_equal(out["coor_xy"].cpu().numpy().reshape(-1, 2),
                          np.stack([x, y], 1).astype(np.float32))
    assert np.all(out["filtered_size"].cpu().numpy() == n_in)


def test_filtered_images_feed_a_batch_plan(lib, torch_cuda):
    """The C-only preparation path: unfiltered images partitioned on the
    device by rc_filter_coarse_images and handed straight to a batch plan
    give the same bits as the host-partitioned inputs, and match float64."""
    from relion_amd import synth
    torch = torch_cuda
    prob = synth.make_problem("c1", seed=11)
    pb, g = prob.particles, prob.geom
    n = g.image_size
    inv = _inverse(synth.filter_order(g)[0])
    fimg = np.concatenate([pb.real[:, inv], pb.imag[:, inv]], axis=1)
    out = lib.filter_images(lib.geometry_view(g.imgX, g.imgY, g.maxR, g.pad), n,
                            _t(torch, fimg), _t(torch, pb.corr[:, inv]))
    assert int(out["filtered_size"][0]) == pb.npix
    projs = [lib.Projector(v) for v in prob.vols]
    host = lib.problem_tensors(prob)
    dev = dict(host, real=out["real"].reshape(-1), imag=out["imag"].reshape(-1),
               corr=out["corr"].reshape(-1), coor_xy=out["coor_xy"],
               diff2=torch.empty_like(host["diff2"]),
               best_diff2=torch.empty_like(host["best_diff2"]),
               best_index=torch.empty_like(host["best_index"]))
    plan = lib.plan_for_problem(prob, projs, "bf16x3")
    plan.execute(host)
    plan.execute(dev)
    torch.cuda.synchronize()
    assert torch.equal(host["diff2"], dev["diff2"])
    assert torch.equal(host["best_index"], dev["best_index"])
    off = lib.output_offsets(prob.trans_offset, prob.n_classes, prob.orientation_num)
    ref = oracle.problem_diff2_expanded(prob, [0])
    got = dev["diff2"][off[0]:off[1]].cpu().numpy().astype(np.float64)
    _check(got, ref[0].ravel(), "bf16x3", "filtered c1 particle 0")
    plan.close()


def test_autotune_choice_reaches_later_plans(lib, torch_cuda):
    """rc_autotune returns one of the listed variants, with a time, and a
    plan created afterwards for the same shape uses it; the process-wide
    override stays automatic."""
    from relion_amd import synth
    prob = synth.make_problem("c1", seed=11)
    projs = [lib.Projector(v) for v in prob.vols]
    t = lib.problem_tensors(prob)
    probe = lib.plan_for_problem(prob, projs, "bf16x3")
    v, ms = probe.autotune(t, iters=2)
    probe.close()
    assert v in [x[0] for x in lib.variants("bf16x3")] and ms > 0
    plan = lib.plan_for_problem(prob, projs, "bf16x3")
    assert plan.tile()[0] == v
    # another shape (fewer particles) keeps the size rule
    sub = synth.make_problem("c1", seed=11, particles=3)
    other = lib.plan_for_problem(sub, projs, "bf16x3")
    assert other.tile()[0] in [x[0] for x in lib.variants("bf16x3")]


@pytest.mark.parametrize("cc", [False, True])
def test_drop_in_non_integer_padding(lib, torch_cuda, cc):
    """The drop-in at padding factor 1.7 (maxR2_padded = int(16^2 1.7^2) =
    739) against the float64 oracle over the unfiltered 32^2 image: pixels
    with (x^2 + y^2) pad^2 just above 739 (739.84 at radius 16) still project
    inside the mask (int(r^2) <= 739), so a search restricted by the filter
    rule alone would drop them; the drop-in's restriction keeps every pixel a
    rotation can bring inside."""
    import dataclasses

    from relion_amd import synth
    torch = torch_cuda
    prob = synth.make_problem("c1", particles=1, seed=11)
    pb, g = prob.particles, prob.geom
    inv = _inverse(synth.filter_order(g)[0])
    vol = dataclasses.replace(prob.vols[0], pad=1.7)
    proj = lib.Projector(vol)
    pk = proj.kernel(g.imgX, g.imgY, g.maxR)
    assert pk.maxR2_padded == 739
    O = prob.orientation_num
    tx, ty = prob.trans_of(0)
    T = tx.size
    re, im, corr = pb.real[0][inv], pb.imag[0][inv], pb.corr[0][inv]
    p = np.arange(g.image_size)
    x, y = p % g.imgX, p // g.imgX
    if cc:
        park = (y > g.maxR) & (y < g.imgY - g.maxR)
        y = np.where((y > g.maxR) & ~park, y - g.imgY, y)
        x = np.where(park, g.maxR, x)
        ref = oracle.cc_f64(vol, pk.maxR2_padded, g.imgX, g.imgY, g.maxR, prob.eulers, tx, ty,
                            re, im, corr).ravel()
    else:
        y = np.where(y > g.maxR, y - g.imgY, y)
        ref = oracle.diff2_f64(vol, pk.maxR2_padded, prob.eulers, tx, ty, re, im, corr,
                               x.astype(np.float64), y.astype(np.float64),
                               n_proj=g.image_size).ravel()
    # pixels beyond the filter rule that the mask still admits for some rotation
    r2 = (x * x + y * y).astype(np.float32) * np.float32(pk.padding_factor) ** 2
    assert np.count_nonzero((r2 > 739) & (r2 < 740)) > 0
    d2 = torch.zeros(O * T, dtype=torch.float32, device="cuda")
    lib.run_diff2_kernel_coarse(pk, _t(torch, tx), _t(torch, ty), _t(torch, corr),
                                _t(torch, re), _t(torch, im), _t(torch, prob.eulers.ravel()), d2,
                                O, T, g.image_size, do_CC=cc)
    torch.cuda.synchronize()
    got = d2.cpu().numpy()
    if cc:
        # the score crosses zero: its error against the score's scale, with
        # the bound of the 32^2 drop-in CC test
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"drop-in pad 1.7 CC: max abs err / max |score| = {err:.3e}")
        assert err <= 1e-5
    else:
        # the bounds of the FP32 form on the 32^2 workload
        st = oracle.error_stats(got, ref)
        print(f"drop-in pad 1.7: max_rel={st['max_rel']:.3e} mean_rel={st['mean_rel']:.3e}")
        assert st["max_rel"] <= 1.2e-6 and st["mean_rel"] <= 2.2e-7, st
    proj.close()
