"""Parity of the HIP paths against the float64 CPU oracle (GPU required).

Every test calls the gfx950 kernels through the C ABI (ctypes) and compares
diff2 element-wise with ``oracle.diff2_f64`` using the reference's error
metrics (check.cuh:96-111).  Tolerances (relative to the float64 value):

Each bound is about 5x the error measured on MI355X for that form and
workload, so a tenfold precision regression fails:

* 32^2 workload (8 particles x 576 x 21): FP32 MFMA max 1.2e-6 / mean
  2.2e-7 (measured 2.4e-7 / 4.3e-8), BF16x3 3.5e-6 / 6e-7 (7e-7 / 1.2e-7),
  FP16 (per-image scaled, one product) 2.5e-4 / 4e-5 (4.7e-5 / 7.4e-6),
  FP32 direct form 7.5e-6 / 2.2e-6 (1.5e-6 / 4.5e-7);
* ragged two-class batch (16^2 images of 32^3 maps at SNR 0.5, more
  cancellation in the score): FP32 7.5e-6 / 1e-6 (1.4e-6 / 2.1e-7), BF16x3
  2.8e-5 / 1e-5 (5.6e-6 / 2.0e-6);
* FP16's arg-min must be a pose whose float64 score is within twice the
  largest FP16 error of the float64 minimum (the agreement rate is printed);
* cross-correlation scores (which cross zero): max |error| <= 7.5e-6 (FP32),
  2e-5 (BF16x3) or 1.5e-3 (FP16) of the largest |score| (measured 1.5e-6,
  4e-6, 2.9e-4).
"""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

TOL = {"fp32": (1.2e-6, 2.2e-7), "fp32_direct": (7.5e-6, 2.2e-6), "bf16x3": (3.5e-6, 6e-7),
       "fp16": (2.5e-4, 4e-5)}
RAGGED_TOL = {"fp32": (7.5e-6, 1e-6), "bf16x3": (2.8e-5, 1e-5)}
CC_TOL = {"fp32": 7.5e-6, "bf16x3": 2e-5, "fp16": 1.5e-3}


@pytest.fixture(scope="module")
def c1():
    from relion_amd import synth
    prob = synth.make_problem("c1", seed=11)
    ref = [oracle.problem_diff2(prob, i) for i in range(prob.n_particles)]
    return prob, ref


@pytest.fixture(scope="module")
def lib(torch_cuda):
    from relion_amd import lib as L
    L.load()
    return L


def _t(torch, a, dt=None):
    x = torch.as_tensor(np.ascontiguousarray(a))
    if dt is not None:
        x = x.to(dt)
    return x.cuda()


def _assert_close_scaled(got, ref, tol, what):
    """Cross-correlation scores cross zero, so their error is measured
    against the score's scale: max |got - ref| <= tol * max |ref|."""
    got = np.asarray(got, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"{what} scaled_max_err={err:.3e}")
    assert err <= tol, (what, err)


def _assert_close(got, ref, form, what, tol=None):
    st = oracle.error_stats(got, ref)
    mx, mean = (tol or TOL)[form]
    print(f"{what} [{form}] max_rel={st['max_rel']:.3e} mean_rel={st['mean_rel']:.3e}")
    assert st["max_rel"] <= mx and st["mean_rel"] <= mean, (what, form, st)


@pytest.mark.parametrize("form", ["fp32", "bf16x3", "fp16", "fp32_direct"])
def test_single_run_filtered(c1, lib, torch_cuda, form):
    """rc_coarse_run on filtered inputs (in-radius first, explicit coords),
    diff2s_dest = diff2s + score with distinct source and destination."""
    torch = torch_cuda
    prob, ref = c1
    pb = prob.particles
    g = prob.geom
    proj = lib.Projector(prob.vols[0])
    pk = proj.kernel(g.imgX, g.imgY, g.maxR)
    O = prob.orientation_num
    eul = _t(torch, prob.eulers.ravel())
    coor = _t(torch, pb.coor_xy.ravel())
    rng = np.random.default_rng(5)
    for i in range(prob.n_particles):
        tx, ty = prob.trans_of(i)
        T = tx.size
        init = rng.uniform(0, 10, O * T).astype(np.float32)
        k = lib.CoarseKernel(T, O, pb.npix, pb.full_npix, form)
        src = _t(torch, init)
        dst = torch.full((O * T,), -1.0, dtype=torch.float32, device="cuda")
        k.run(eul, _t(torch, tx), _t(torch, ty), _t(torch, pb.real[i]), _t(torch, pb.imag[i]), pk,
              _t(torch, pb.corr[i]), src, dst, coor_xy=coor)
        torch.cuda.synchronize()
        got = dst.cpu().numpy().astype(np.float64)
        _assert_close(got, ref[i].ravel() + init, form, f"particle {i}")
        assert np.array_equal(src.cpu().numpy(), init)


def test_run_diff2_kernel_coarse_unfiltered(c1, lib, torch_cuda):
    """Mirror of runDiff2KernelCoarse: unfiltered image, coordinates from the
    pixel index, diff2s accumulated in place."""
    from relion_amd import synth
    torch = torch_cuda
    prob, ref = c1
    pb = prob.particles
    g = prob.geom
    order, _ = synth.filter_order(g)
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size)
    proj = lib.Projector(prob.vols[0])
    pk = proj.kernel(g.imgX, g.imgY, g.maxR)
    O = prob.orientation_num
    eul = _t(torch, prob.eulers.ravel())
    for i in (0, 3):
        tx, ty = prob.trans_of(i)
        T = tx.size
        init = np.full(O * T, 2.5, np.float32)
        d2 = _t(torch, init)
        lib.run_diff2_kernel_coarse(pk, _t(torch, tx), _t(torch, ty), _t(torch, pb.corr[i][inv]),
                                    _t(torch, pb.real[i][inv]), _t(torch, pb.imag[i][inv]), eul,
                                    d2, O, T, g.image_size)
        torch.cuda.synchronize()
        _assert_close(d2.cpu().numpy(), ref[i].ravel() + 2.5, "fp32", f"unfiltered {i}")


def test_run_diff2_kernel_coarse_cc_mode(c1, lib, torch_cuda):
    """do_CC: -sum corr Re(P conj Xs) / sqrt(sum corr |P|^2), written over
    diff2s (diff2.cuh:691-809), against the float64 CC oracle."""
    from relion_amd import synth
    torch = torch_cuda
    prob, _ = c1
    pb = prob.particles
    g = prob.geom
    order, _ = synth.filter_order(g)
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size)
    proj = lib.Projector(prob.vols[0])
    pk = proj.kernel(g.imgX, g.imgY, g.maxR)
    O = prob.orientation_num
    eul = _t(torch, prob.eulers.ravel())
    for i in (1, 6):
        tx, ty = prob.trans_of(i)
        T = tx.size
        d2 = torch.full((O * T,), 123.0, dtype=torch.float32, device="cuda")
        lib.run_diff2_kernel_coarse(pk, _t(torch, tx), _t(torch, ty), _t(torch, pb.corr[i][inv]),
                                    _t(torch, pb.real[i][inv]), _t(torch, pb.imag[i][inv]), eul,
                                    d2, O, T, g.image_size, do_CC=True)
        torch.cuda.synchronize()
        ref = oracle.cc_f64(prob.vols[0], g.maxR2_padded, g.imgX, g.imgY, g.maxR, prob.eulers,
                            tx, ty, pb.real[i][inv], pb.imag[i][inv], pb.corr[i][inv])
        _assert_close_scaled(d2.cpu().numpy(), ref.ravel(), 1e-5, f"cc {i}")


def test_run_diff2_kernel_coarse_rejects_3d_data(c1, lib, torch_cuda):
    torch = torch_cuda
    prob, _ = c1
    g = prob.geom
    proj = lib.Projector(prob.vols[0])
    pk = proj.kernel(g.imgX, g.imgY, g.maxR)
    x = torch.zeros(4096, device="cuda")
    with pytest.raises(lib.RcError) as e:
        lib.run_diff2_kernel_coarse(pk, x, x, x, x, x, x, x, 1, 1, 16, data_is_3D=True)
    assert e.value.status == 2


def _wide_256(lib, form):
    """Indices of the 256 x 256 kernels (the wide, LDS-staged family)."""
    return [v for v, tm, tn in lib.variants(form) if tm == 256 and tn == 256]


def _wide_512(lib, form):
    """Indices of the 128 x 512 wide kernels (the large-batch default)."""
    return [v for v, tm, tn in lib.variants(form)
            if tm == 128 and tn == 512 and lib.variant_name(form, v).startswith("cross_kernel_wide")]


def _direct(lib, form):
    """Indices of the direct kernels (projections loaded into registers)."""
    return [v for v, _, _ in lib.variants(form)
            if lib.variant_name(form, v).startswith("cross_kernel_direct")]


@pytest.mark.parametrize("kernel", ["default", "wide", "wide512", "direct", "direct_pre"])
@pytest.mark.parametrize("form", ["fp32", "bf16x3", "fp16"])
def test_batched_cc(c1, lib, torch_cuda, form, kernel):
    """Batched cross-correlation scores; the per-particle minimum is the best
    correlation."""
    from relion_amd import synth
    torch = torch_cuda
    prob, _ = c1
    pb = prob.particles
    g = prob.geom
    order, _ = synth.filter_order(g)
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size)
    projs = [lib.Projector(v) for v in prob.vols]
    try:
        if kernel == "wide":
            lib.set_variant_override(_wide_256(lib, form)[0])
        elif kernel == "wide512":
            lib.set_variant_override(_wide_512(lib, form)[-1])
        elif kernel == "direct":
            lib.set_variant_override(_direct(lib, form)[0])
        elif kernel == "direct_pre":
            lib.set_variant_override(_direct(lib, form)[1])
        plan = lib.plan_for_problem(prob, projs, form, cc=True)
        t = lib.problem_tensors(prob)
        plan.execute(t)
        torch.cuda.synchronize()
    finally:
        lib.set_variant_override(-1)
    d = t["diff2"].cpu().numpy().astype(np.float64)
    off = lib.output_offsets(prob.trans_offset, 1, prob.orientation_num)
    bidx = t["best_index"].cpu().numpy()
    for i in (0, 4):
        tx, ty = prob.trans_of(i)
        ref = oracle.cc_f64(prob.vols[0], g.maxR2_padded, g.imgX, g.imgY, g.maxR, prob.eulers,
                            tx, ty, pb.real[i][inv], pb.imag[i][inv], pb.corr[i][inv])
        _assert_close_scaled(d[off[i]:off[i + 1]], ref.ravel(), CC_TOL[form],
                             f"batched cc {i} ({kernel})")
        if form != "fp16":
            assert bidx[i] == oracle.argmin_flat(ref)


@pytest.mark.parametrize("form", ["fp32", "bf16x3", "fp16"])
def test_batched_c1(c1, lib, torch_cuda, form):
    torch = torch_cuda
    prob, ref = c1
    projs = [lib.Projector(v) for v in prob.vols]
    plan = lib.plan_for_problem(prob, projs, form)
    t = lib.problem_tensors(prob)
    plan.execute(t)
    torch.cuda.synchronize()
    d = t["diff2"].cpu().numpy().astype(np.float64)
    off = lib.output_offsets(prob.trans_offset, prob.n_classes, prob.orientation_num)
    best = t["best_diff2"].cpu().numpy()
    bidx = t["best_index"].cpu().numpy()
    agree = 0
    for i in range(prob.n_particles):
        got = d[off[i]:off[i + 1]]
        _assert_close(got, ref[i].ravel(), form, f"batched particle {i}")
        r = ref[i].ravel()
        k = oracle.argmin_flat(ref[i])
        agree += int(bidx[i] == k)
        if form != "fp16":
            assert bidx[i] == k, (i, bidx[i], k)
        else:
            # a near-optimal pose: its float64 score is within twice the
            # largest FP16 error of the float64 minimum
            assert r[bidx[i]] - r[k] <= 2.0 * np.abs(got - r).max(), (i, bidx[i], k)
        assert abs(best[i] - got[bidx[i]]) <= 1e-6 * abs(best[i])
        assert got[bidx[i]] == got.min()
    print(f"[{form}] arg-min agreement with float64: {agree}/{prob.n_particles}")


def _ragged_problem():
    """Two classes, uneven translation counts (3..37 shifts), 5 particles."""
    from relion_amd import synth
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
    return synth.CoarseProblem("ragged", vols, g, eul, np.asarray(offs, np.int32),
                               np.concatenate(txs), np.concatenate(tys), pb, 32)


def test_batched_ragged_multiclass(lib, torch_cuda):
    """Two classes, uneven translation counts (3..37 shifts), 5 particles."""
    torch = torch_cuda
    prob = _ragged_problem()
    vols = prob.vols
    off = lib.output_offsets(prob.trans_offset, 2, prob.orientation_num)
    refs = [np.stack([oracle.problem_diff2(prob, i, cls=c) for c in range(2)])
            for i in range(prob.n_particles)]
    projs = [lib.Projector(v) for v in vols]
    try:
        for form in ("fp32", "bf16x3"):
            for v, tm, tn in lib.variants(form):
                lib.set_variant_override(v)
                plan = lib.plan_for_problem(prob, projs, form)
                t = lib.problem_tensors(prob)
                plan.execute(t)
                torch.cuda.synchronize()
                d = t["diff2"].cpu().numpy().astype(np.float64)
                bidx = t["best_index"].cpu().numpy()
                for i in range(prob.n_particles):
                    _assert_close(d[off[i]:off[i + 1]], refs[i].ravel(), form,
                                  f"ragged variant {v} ({tm}x{tn}) particle {i}", RAGGED_TOL)
                    assert bidx[i] == oracle.argmin_flat(refs[i]), (form, v, i)
                plan.close()
    finally:
        lib.set_variant_override(-1)


def test_batched_diff2_init_all_variants(lib, torch_cuda):
    """d_diff2_init is added to every score in float64 before the one rounding
    to float32, and the per-particle minimum is taken over the sums: for
    every kernel and tile shape on the ragged two-class batch, against the
    same plan run without it (sum within one rounding of each, minimum and
    its index on the sums)."""
    torch = torch_cuda
    prob = _ragged_problem()
    off = lib.output_offsets(prob.trans_offset, 2, prob.orientation_num)
    projs = [lib.Projector(v) for v in prob.vols]
    rng = np.random.default_rng(23)
    try:
        for form in ("fp32", "bf16x3", "fp16"):
            for v, tm, tn in lib.variants(form):
                lib.set_variant_override(v)
                plan = lib.plan_for_problem(prob, projs, form)
                t = lib.problem_tensors(prob)
                plan.execute(t)
                torch.cuda.synchronize()
                d0 = t["diff2"].cpu().numpy().astype(np.float64)
                init = (rng.standard_normal(d0.size) * d0.std()).astype(np.float32)
                t["diff2_init"] = torch.as_tensor(init).cuda()
                t["diff2"].fill_(float("nan"))
                plan.execute(t)
                torch.cuda.synchronize()
                d1 = t["diff2"].cpu().numpy().astype(np.float64)
                err = np.abs(d1 - (d0 + init))
                bound = 1.5e-7 * (np.abs(d0) + np.abs(d1)) + 1e-30
                assert np.all(err <= bound), (form, v, float((err / bound).max()))
                best = t["best_diff2"].cpu().numpy()
                bidx = t["best_index"].cpu().numpy()
                for i in range(prob.n_particles):
                    got = d1[off[i]:off[i + 1]]
                    assert got[bidx[i]] == got.min(), (form, v, i)
                    assert abs(best[i] - got[bidx[i]]) <= 1e-6 * abs(best[i]), (form, v, i)
                plan.close()
    finally:
        lib.set_variant_override(-1)


def test_batched_variants_agree(c1, lib, torch_cuda):
    """Every compiled kernel and tile shape meets its form's tolerance."""
    torch = torch_cuda
    prob, ref = c1
    projs = [lib.Projector(v) for v in prob.vols]
    off = lib.output_offsets(prob.trans_offset, 1, prob.orientation_num)
    try:
        for form in ("fp32", "bf16x3", "fp16"):
            for v, tm, tn in lib.variants(form):
                lib.set_variant_override(v)
                plan = lib.plan_for_problem(prob, projs, form)
                t = lib.problem_tensors(prob)
                plan.execute(t)
                torch.cuda.synchronize()
                d = t["diff2"].cpu().numpy().astype(np.float64)
                for i in (0, 7):
                    _assert_close(d[off[i]:off[i + 1]], ref[i].ravel(), form,
                                  f"variant {v} ({tm}x{tn}) particle {i}")
    finally:
        lib.set_variant_override(-1)


@pytest.mark.parametrize("kernel", ["default", "wide"])
def test_plan_graph_replay_matches_eager(c1, lib, torch_cuda, kernel):
    """A captured plan replayed several times, with eager work and host
    synchronisations between replays, reproduces the eager result exactly."""
    torch = torch_cuda
    prob, ref = c1
    projs = [lib.Projector(v) for v in prob.vols]
    try:
        if kernel == "wide":
            lib.set_variant_override(_wide_256(lib, "bf16x3")[-1])
        plan = lib.plan_for_problem(prob, projs, "bf16x3")
    finally:
        lib.set_variant_override(-1)
    t = lib.problem_tensors(prob)
    plan.execute(t)
    torch.cuda.synchronize()
    eager = t["diff2"].clone()
    eager_best = t["best_index"].clone()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        plan.capture(t, s)
    torch.cuda.synchronize()
    for k in range(3):
        t["diff2"].fill_(-1.0)
        t["best_index"].fill_(-1)
        torch.cuda.synchronize()
        plan.replay(s)
        s.synchronize()
        assert torch.equal(t["diff2"], eager), k
        assert torch.equal(t["best_index"], eager_best), k
        plan.execute(t)          # eager work between replays
        torch.cuda.synchronize()


@pytest.mark.parametrize("kernel", ["default", "wide"])
def test_batched_deterministic_alone_and_in_batch(c1, lib, torch_cuda, kernel):
    """Each particle's diff2 is bit-identical whether it runs alone or inside
    a larger batch (no atomics, no split-K)."""
    if kernel == "wide":
        lib.set_variant_override(_wide_256(lib, "bf16x3")[1])
        try:
            _deterministic_alone_and_in_batch(c1, lib, torch_cuda)
        finally:
            lib.set_variant_override(-1)
    else:
        _deterministic_alone_and_in_batch(c1, lib, torch_cuda)


def _deterministic_alone_and_in_batch(c1, lib, torch_cuda):
    from relion_amd import synth
    torch = torch_cuda
    prob, _ = c1
    projs = [lib.Projector(v) for v in prob.vols]
    full = lib.plan_for_problem(prob, projs, "bf16x3")
    t = lib.problem_tensors(prob)
    full.execute(t)
    torch.cuda.synchronize()
    d_full = t["diff2"].cpu().numpy()
    off = lib.output_offsets(prob.trans_offset, 1, prob.orientation_num)
    pb = prob.particles
    for i in (0, 5):
        one = synth.ParticleBatch(pb.real[i:i + 1], pb.imag[i:i + 1], pb.corr[i:i + 1],
                                  pb.coor_xy, pb.npix, pb.full_npix, pb.true_rot[i:i + 1],
                                  pb.true_shift[i:i + 1], pb.classes[i:i + 1], pb.snr)
        tx, ty = prob.trans_of(i)
        p1 = synth.CoarseProblem("one", prob.vols, prob.geom, prob.eulers,
                                 np.array([0, tx.size], np.int32), tx, ty, one, prob.box)
        plan = lib.plan_for_problem(p1, projs, "bf16x3")
        t1 = lib.problem_tensors(p1)
        plan.execute(t1)
        torch.cuda.synchronize()
        assert np.array_equal(t1["diff2"].cpu().numpy(), d_full[off[i]:off[i + 1]])


def _plain_and_precomputed(lib, form):
    """(plain-schedule variant, translation-operand-precomputed variant) pairs
    of the same 512-thread wide tile shape."""
    names = {v: lib.variant_name(form, v) for v, _, _ in lib.variants(form)}
    pairs = []
    for v, n in names.items():
        if "translation operand precomputed" not in n or not n.startswith("cross_kernel_wide"):
            continue
        shape = n.split(",")[0]                 # "cross_kernel_wide<...> 128x512"
        plain = [u for u, m in names.items()
                 if m.startswith(shape + ", 512 threads") and m.count(",") == 2]
        assert len(plain) == 1, (n, plain)
        pairs.append((plain[0], v))
    return pairs


@pytest.mark.parametrize("form", ["fp32", "bf16x3", "fp16"])
def test_precomputed_operand_is_bit_identical(c1, lib, torch_cuda, form):
    """The precomputed translation operand is built by the same arithmetic
    as the in-kernel build and the MFMAs run in the same order, so each
    such variant reproduces the plain variant of its tile shape bit for bit
    (the 32^2 workload and the ragged two-class batch)."""
    torch = torch_cuda
    pairs = _plain_and_precomputed(lib, form)
    assert len(pairs) >= 2
    for prob in (c1[0], _ragged_problem()):
        projs = [lib.Projector(v) for v in prob.vols]
        try:
            for plain, pre in pairs:
                got = []
                for v in (plain, pre):
                    lib.set_variant_override(v)
                    plan = lib.plan_for_problem(prob, projs, form)
                    t = lib.problem_tensors(prob)
                    plan.execute(t)
                    torch.cuda.synchronize()
                    got.append((t["diff2"].cpu().numpy(), t["best_index"].cpu().numpy()))
                    plan.close()
                assert np.array_equal(got[0][0], got[1][0]), (form, plain, pre)
                assert np.array_equal(got[0][1], got[1][1]), (form, plain, pre)
        finally:
            lib.set_variant_override(-1)


@pytest.mark.parametrize("form", ["fp32", "bf16x3", "fp16"])
def test_direct_kernel_is_bit_identical(c1, lib, torch_cuda, form):
    """The direct kernels (projections loaded straight into registers, one
    wave per SIMD; the translation operand built in the kernel, or
    precomputed and copied) build the same operands and run every
    accumulator's MFMAs in the same order as the wide 128 x 512 kernels, so
    they reproduce them bit for bit: the 32^2 workload (a partial last column
    tile of 64 orientations) and the ragged two-class batch (a partial last
    row tile)."""
    torch = torch_cuda
    names = {v: lib.variant_name(form, v) for v, _, _ in lib.variants(form)}
    plain = [v for v, n in names.items()
             if n.startswith("cross_kernel_wide") and "128x512, 512 threads" in n
             and n.count(",") == 2]
    direct = _direct(lib, form)
    assert len(plain) == 1 and len(direct) == 2, names
    assert "translation operand precomputed" in names[direct[1]]
    for prob in (c1[0], _ragged_problem()):
        projs = [lib.Projector(v) for v in prob.vols]
        try:
            got = []
            for v in (plain[0], *direct):
                lib.set_variant_override(v)
                plan = lib.plan_for_problem(prob, projs, form)
                t = lib.problem_tensors(prob)
                plan.execute(t)
                torch.cuda.synchronize()
                got.append((t["diff2"].cpu().numpy(), t["best_index"].cpu().numpy(),
                            t["best_diff2"].cpu().numpy()))
                plan.close()
            for other in got[1:]:
                for a, b in zip(got[0], other):
                    assert np.array_equal(a, b), (form, plain, direct)
        finally:
            lib.set_variant_override(-1)
