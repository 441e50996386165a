"""Pins the CPU oracle against facts the mathematics fixes (CPU only).

The reference ships no golden vectors for this path (its inputs are
real-data dumps that are not in the repository), so each part of the
oracle is checked against an independent fact chosen so that a plausible
mistake fails at least one test:

projection   Fourier-slice theorem at identity and 90-degree rotations
             (volume layout, initY/initZ, padding, e-matrix rows), Hermitian
             pairing against a full complex FFT (flip + conjugation), exact
             reproduction of linear fields (trilinear weights), brute-force
             clamp-to-edge, the int-truncated r^2 boundary.
phase        shift theorem: an integer shift by e^{i phi} equals the FFT of
             the circularly shifted image (sign and axis of the phase).
score        brute-force float64 triple loop, the expanded form
             |P|^2 + |X|^2 - 2 Re(P conj Xs), linearity in the initial
             value, pixel-permutation invariance (filtered vs unfiltered
             order), an exact zero for a noise-free particle, recovery of
             the generating pose from noisy particles.
"""
import math

import numpy as np
import pytest

from oracle import oracle
from relion_amd import synth


def _small_volume(n=16, seed=2):
    return synth.fourier_volume(synth.blob_map(n, n_blobs=6, seed=seed, sigma_range=(1.0, 2.5)))


def _full_fft(n=16, seed=2, pad=2):
    """Full complex FFT of the same padded, centred map (scaled as the volume)."""
    m = synth.blob_map(n, n_blobs=6, seed=seed, sigma_range=(1.0, 2.5))
    p = pad * n
    big = np.zeros((p, p, p))
    o = (p - n) // 2
    big[o:o + n, o:o + n, o:o + n] = m
    return np.fft.fftn(np.fft.ifftshift(big)) / float(n) ** 1.5, big


def test_fourier_slice_identity():
    n = 16
    v = _small_volume(n)
    F, big = _full_fft(n)
    p = 2 * n
    g = synth.CoarseGeometry(imgY=n, maxR=n // 2)
    x, y = synth.pixel_coords(g)
    eye = np.eye(3, dtype=np.float32).ravel()
    got = oracle.project(v, g.maxR2_padded, eye, x, y)
    # central slice = 2D FFT of the z-projection
    slice2d = np.fft.fft2(np.fft.ifftshift(big.sum(axis=0))) / float(n) ** 1.5
    want = slice2d[(2 * y) % p, (2 * x) % p]
    inside = (x * x + y * y) * 4 <= g.maxR2_padded
    want = np.where(inside, want, 0)
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-6 * np.abs(want).max())
    assert np.abs(got[inside]).max() > 0


@pytest.mark.parametrize("axis,angle", [("z", 90), ("x", 90), ("y", 90), ("z", 180),
                                        ("x", 270)])
def test_hermitian_pairing_right_angles(axis, angle):
    """At right-angle rotations with pad=2 every sample hits a voxel centre,
    so the half-volume projection (with its Hermitian flip for negative
    rotated x) must equal the full complex FFT looked up at 2 M (x, y, 0)."""
    n = 16
    v = _small_volume(n)
    F, _ = _full_fft(n)
    p = 2 * n
    c, s = int(round(math.cos(math.radians(angle)))), int(round(math.sin(math.radians(angle))))
    M = {"z": [[c, -s, 0], [s, c, 0], [0, 0, 1]],
         "x": [[1, 0, 0], [0, c, -s], [0, s, c]],
         "y": [[c, 0, s], [0, 1, 0], [-s, 0, c]]}[axis]
    M = np.asarray(M, dtype=np.float32)
    g = synth.CoarseGeometry(imgY=n, maxR=n // 2)
    x, y = synth.pixel_coords(g)
    got = oracle.project(v, g.maxR2_padded, M.ravel(), x, y)
    k = 2 * (M.astype(np.int64) @ np.stack([x, y, np.zeros_like(x)]))
    want = F[k[2] % p, k[1] % p, k[0] % p]
    inside = (x * x + y * y) * 4 <= g.maxR2_padded
    want = np.where(inside, want, 0)
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-6 * np.abs(want).max())
    # a transposed matrix gives a different answer for the non-symmetric x rotation
    if axis == "x":
        bad = oracle.project(v, g.maxR2_padded, M.T.copy().ravel(), x, y)
        assert np.abs(bad - want).max() > 1e-3 * np.abs(want).max()


def test_trilinear_reproduces_linear_fields():
    rng = np.random.default_rng(0)
    mx, my, mz = 20, 41, 41
    zz, yy, xx = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    a = rng.normal(size=4)
    b = rng.normal(size=4)
    re = (a[0] + a[1] * xx + a[2] * yy + a[3] * zz).astype(np.float32).ravel()
    im = (b[0] + b[1] * xx + b[2] * yy + b[3] * zz).astype(np.float32).ravel()
    v = synth.ProjectorVolume(re, im, mx, my, mz, -20, -20, 8, 2.0)
    for _ in range(5):
        m9 = synth._random_rotation(rng).astype(np.float32).ravel()
        x = rng.integers(0, 8, 30).astype(np.float64)
        y = rng.integers(-8, 8, 30).astype(np.float64)
        got = oracle.project(v, 10 ** 6, m9, x, y)
        m = m9.astype(np.float64).reshape(3, 3)
        k = 2.0 * (m @ np.stack([x, y, np.zeros_like(x)]))
        neg = k[0] < 0
        k = np.where(neg, -k, k)
        X, Y, Z = k[0], k[1] + 20, k[2] + 20
        wre = a[0] + a[1] * X + a[2] * Y + a[3] * Z
        wim = b[0] + b[1] * X + b[2] * Y + b[3] * Z
        want = wre + 1j * np.where(neg, -wim, wim)
        # voxel values are stored as float32: 1 ulp of ~60
        np.testing.assert_allclose(got, want, rtol=3e-7, atol=1e-6)


def test_clamp_to_edge_brute_force():
    rng = np.random.default_rng(1)
    mx, my, mz = 4, 5, 6
    re = rng.normal(size=mx * my * mz).astype(np.float32)
    im = rng.normal(size=mx * my * mz).astype(np.float32)
    v = synth.ProjectorVolume(re, im, mx, my, mz, -2, -3, 100, 1.0)
    vol = (re.astype(np.float64) + 1j * im).reshape(mz, my, mx)

    def brute(X, Y, Z):
        acc = 0j
        x0, y0, z0 = math.floor(X), math.floor(Y), math.floor(Z)
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    w = (X - x0 if dx else 1 - (X - x0)) * (Y - y0 if dy else 1 - (Y - y0)) * \
                        (Z - z0 if dz else 1 - (Z - z0))
                    acc += w * vol[min(max(z0 + dz, 0), mz - 1), min(max(y0 + dy, 0), my - 1),
                                   min(max(x0 + dx, 0), mx - 1)]
        return acc

    eye = np.eye(3, dtype=np.float32).ravel()
    for _ in range(200):
        # identity orientation: (xp, yp, zp) = (x, y, 0) with pad 1; x >= 0
        x = float(rng.uniform(0, 6))
        y = float(rng.uniform(-6, 6))
        got = oracle.project(v, 10 ** 6, eye, [x], [y])[0]
        want = brute(x, y + 2, 0 + 3)   # initY = -2, initZ = -3
        assert abs(got - want) < 1e-9 * (1 + abs(want))


def test_r2_boundary_is_int_truncated():
    v = _small_volume(16)
    eye = np.eye(3, dtype=np.float32).ravel()
    maxR2 = 4 * 8 * 8
    # (x, y) = (8, 0): 4*64 = maxR2 -> inside; (8, 1): 4*65 -> outside
    got = oracle.project(v, maxR2, eye, [8, 8, 0], [0, 1, 8])
    assert got[0] != 0 and got[1] == 0 and got[2] != 0
    # int truncation: r2 = 256.75 truncates to 256 <= 256
    s = np.float32(math.sqrt(256.75) / 2.0 / 8.0)
    got = oracle.project(v, maxR2, np.array([s, 0, 0, 0, s, 0, 0, 0, 1], np.float32), [8], [0])
    assert got[0] != 0


def test_shift_theorem():
    rng = np.random.default_rng(3)
    n = 32
    img = rng.normal(size=(n, n))
    X = np.fft.fft2(img)
    g = synth.CoarseGeometry(imgY=n, maxR=n // 2)
    x, y = synth.pixel_coords(g)
    Xh = X[y % n, x % n]
    for dx, dy in [(1, 0), (0, 1), (-3, 2), (5, -4)]:
        tx, ty = 2 * math.pi * dx / n, 2 * math.pi * dy / n
        got = oracle.translate(x, y, np.float32(tx), np.float32(ty), Xh.real, Xh.imag)
        want = np.fft.fft2(np.roll(img, (-dy, -dx), axis=(0, 1)))[y % n, x % n]
        # tx, ty go through float32: phase error ~ |k| * 1e-7
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-5 * np.abs(want).max())
        assert np.allclose(np.abs(got), np.abs(Xh.astype(np.complex64)), rtol=1e-6)


def _tiny_problem(seed=4):
    rng = np.random.default_rng(seed)
    v = _small_volume(16, seed=seed)
    g = synth.CoarseGeometry(imgY=16, maxR=8)
    order, npix = synth.filter_order(g)
    x, y = synth.pixel_coords(g)
    x, y = x[order].astype(np.float64), y[order].astype(np.float64)
    n = x.size
    O, T = 5, 4
    eul = np.stack([synth._random_rotation(rng).ravel() for _ in range(O)]).astype(np.float32)
    tx = rng.uniform(-0.5, 0.5, T).astype(np.float32)
    ty = rng.uniform(-0.5, 0.5, T).astype(np.float32)
    re = rng.normal(size=n).astype(np.float32)
    im = rng.normal(size=n).astype(np.float32)
    w = rng.uniform(0, 2, n).astype(np.float32)
    return v, g, x, y, npix, eul, tx, ty, re, im, w


def test_score_brute_force_and_expansion():
    v, g, x, y, npix, eul, tx, ty, re, im, w = _tiny_problem()
    got = oracle.diff2_f64(v, g.maxR2_padded, eul, tx, ty, re, im, w, x, y, n_proj=npix)
    O, T = eul.shape[0], tx.size
    brute = np.zeros((O, T))
    expanded = np.zeros((O, T))
    for o in range(O):
        P = oracle.project(v, g.maxR2_padded, eul[o], x, y)
        P[npix:] = 0
        for t in range(T):
            Xs = oracle.translate(x, y, tx[t], ty[t], re, im)
            brute[o, t] = np.sum(w / 2.0 * np.abs(P - Xs) ** 2)
            expanded[o, t] = (np.sum(w / 2.0 * np.abs(P) ** 2) +
                              np.sum(w / 2.0 * (re.astype(np.float64) ** 2 + im.astype(np.float64) ** 2))
                              - np.sum(w * np.real(P * np.conj(Xs))))
    np.testing.assert_allclose(got, brute, rtol=1e-12)
    np.testing.assert_allclose(got, expanded, rtol=1e-10)
    # the scores depend on orientation and translation (no degenerate input)
    assert got.std() > 1e-3 * got.mean()


def test_score_linear_in_init_and_permutation_invariant():
    v, g, x, y, npix, eul, tx, ty, re, im, w = _tiny_problem(7)
    base = oracle.diff2_f64(v, g.maxR2_padded, eul, tx, ty, re, im, w, x, y, n_proj=npix)
    init = np.random.default_rng(0).normal(size=base.size)
    with_init = oracle.diff2_f64(v, g.maxR2_padded, eul, tx, ty, re, im, w, x, y, n_proj=npix,
                                 init=init)
    np.testing.assert_allclose(with_init, base + init.reshape(base.shape), rtol=1e-13)
    # unfiltered order, all pixels projected (r^2 mask zeroes the tail)
    order, _ = synth.filter_order(g)
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size)
    unf = oracle.diff2_f64(v, g.maxR2_padded, eul, tx, ty, re[inv], im[inv], w[inv], x[inv],
                           y[inv])
    np.testing.assert_allclose(unf, base, rtol=1e-12)
    perm = np.random.default_rng(1).permutation(npix)
    full_perm = np.concatenate([perm, np.arange(npix, x.size)])
    p2 = oracle.diff2_f64(v, g.maxR2_padded, eul, tx, ty, re[full_perm], im[full_perm],
                          w[full_perm], x[full_perm], y[full_perm], n_proj=npix)
    np.testing.assert_allclose(p2, base, rtol=1e-12)


def test_noise_free_particle_scores_zero_at_truth():
    prob = synth.make_problem("mini", seed=21)
    v = prob.vols[0]
    pb = prob.particles
    x, y = pb.coor_xy[:, 0].astype(np.float64), pb.coor_xy[:, 1].astype(np.float64)
    o_star = 17
    P = oracle.project(v, prob.geom.maxR2_padded, prob.eulers[o_star], x, y)
    P[pb.npix:] = 0
    re, im = P.real.astype(np.float32), P.imag.astype(np.float32)
    w = np.ones(x.size, np.float32)
    tx, ty, shifts = synth.translation_set(2, prob.box)
    t0 = int(np.nonzero((shifts[:, 0] == 0) & (shifts[:, 1] == 0))[0][0])
    d = oracle.diff2_f64(v, prob.geom.maxR2_padded, prob.eulers, tx, ty, re, im, w, x, y,
                         n_proj=pb.npix)
    assert d[o_star, t0] < 1e-9 * d.mean()
    assert oracle.argmin_flat(d) == o_star * tx.size + t0


def test_pose_recovery_on_noisy_c1():
    """Particles made at a grid pose (orientation o*, integer shift t*) with
    CTF and noise at SNR 2: the float64 argmin is exactly (o*, t*)."""
    prob = synth.make_problem("c1", seed=5, particles=1)
    g = prob.geom
    v = prob.vols[0]
    order, npix = synth.filter_order(g)
    x, y = synth.pixel_coords(g)
    x, y = x[order].astype(np.float64), y[order].astype(np.float64)
    tx, ty, shifts = synth.translation_set(5, prob.box)
    rng = np.random.default_rng(12)
    s2 = (x * x + y * y) / (prob.box * 1.5) ** 2
    for _ in range(6):
        o_star = int(rng.integers(prob.orientation_num))
        t_star = int(rng.integers(tx.size))
        p = synth.project_np(v, prob.eulers[o_star], x, y, g.maxR2_padded)
        ctf = synth._ctf(s2, rng.uniform(8000, 20000))
        sig = ctf * p * np.exp(-1j * (x * float(tx[t_star]) + y * float(ty[t_star])))
        sigma2 = np.mean(np.abs(sig[:npix]) ** 2) / 2.0
        obs = sig + rng.normal(scale=math.sqrt(sigma2 / 2), size=x.size) * (1 + 1j)
        ok = np.abs(ctf) >= 0.05
        xt = np.where(ok, obs / np.where(ok, ctf, 1), 0)
        w = np.where(ok, ctf * ctf / sigma2, 0)
        d = oracle.diff2_f64(v, g.maxR2_padded, prob.eulers, tx, ty, xt.real, xt.imag, w, x, y,
                             n_proj=npix)
        assert oracle.argmin_flat(d) == o_star * tx.size + t_star


def test_f32_port_matches_f64():
    prob = synth.make_problem("c1", seed=8, particles=2)
    pb = prob.particles
    for i in range(2):
        tx, ty = prob.trans_of(i)
        ref = oracle.problem_diff2(prob, i)
        got, used = oracle.diff2_f32(prob.vols[0], prob.geom.maxR2_padded, prob.eulers, tx, ty,
                                     pb.real[i], pb.imag[i], pb.corr[i], pb.coor_xy, pb.npix)
        st = oracle.error_stats(got, ref)
        assert used >= 1
        assert st["max_rel"] < 2e-5 and st["mean_rel"] < 1e-6, st


def test_error_stats_follow_reference_metrics():
    st = oracle.error_stats([1.0, 2.0, 0.0, 5.0], [1.0, 1.0, 1e-13, 4.0])
    assert st["count"] == 4 and st["count_rel"] == 3
    assert st["max_abs"] == 1.0
    assert abs(st["mean_rel"] - (0 + 1.0 + 0.25) / 3) < 1e-15


def test_cc_score_against_diff2_expansion():
    """Cross-correlation score (diff2.cuh:691-809) pinned by the expanded
    diff2: sum w Re(P conj Xs) = sum w/2 |P|^2 + sum w/2 |X|^2 - diff2."""
    v, g, x, y, npix, eul, tx, ty, re, im, w = _tiny_problem(9)
    order, _ = synth.filter_order(g)
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size)
    xu, yu, reu, imu, wu = x[inv], y[inv], re[inv], im[inv], w[inv]
    cc = oracle.cc_f64(v, g.maxR2_padded, g.imgX, g.imgY, g.maxR, eul, tx, ty, reu, imu, wu)
    d2 = oracle.diff2_f64(v, g.maxR2_padded, eul, tx, ty, reu, imu, wu, xu, yu)
    xn = np.sum(wu / 2.0 * (reu.astype(np.float64) ** 2 + imu.astype(np.float64) ** 2))
    for o in range(eul.shape[0]):
        P = oracle.project(v, g.maxR2_padded, eul[o], xu, yu)
        pn = np.sum(wu / 2.0 * np.abs(P) ** 2)
        want = -(pn + xn - d2[o]) / np.sqrt(2.0 * pn)
        np.testing.assert_allclose(cc[o], want, rtol=1e-9, atol=1e-12 * np.abs(want).max())


def test_expanded_form_equals_direct_form():
    """The expanded-form evaluator (projections once, float64 matrix products)
    used for the full-size GPU checks gives the direct form's values: every
    C1 particle, and a ragged two-class problem with an out-of-radius tail."""
    prob = synth.make_problem("c1", seed=17)
    d = oracle.problem_diff2_expanded(prob, range(prob.n_particles))
    for i in range(prob.n_particles):
        ref = oracle.problem_diff2(prob, i)
        assert np.abs(d[i] - ref).max() <= 1e-12 * np.abs(ref).max()
        assert oracle.argmin_flat(d[i]) == oracle.argmin_flat(ref)
    vols = [synth.fourier_volume(synth.blob_map(32, seed=s)) for s in (3, 4)]
    g = synth.CoarseGeometry(imgY=16, maxR=6)            # many pixels outside the radius
    eul = synth.orientation_set(0, 8)
    pb = synth.make_particles(vols, g, 3, 32, 0.5, seed=9, shift_range=1.0)
    assert pb.full_npix > pb.npix
    txs, tys, offs = [], [], [0]
    for i, r2 in enumerate((1, 13, 4)):
        tx, ty, _ = synth.translation_set(r2, 32, centre=(0.5 * i, -0.25 * i))
        txs.append(tx)
        tys.append(ty)
        offs.append(offs[-1] + tx.size)
    p2 = synth.CoarseProblem("ragged", vols, g, eul, np.asarray(offs, np.int32),
                             np.concatenate(txs), np.concatenate(tys), pb, 32)
    for c in range(2):
        d = oracle.problem_diff2_expanded(p2, range(3), cls=c)
        for i in range(3):
            ref = oracle.problem_diff2(p2, i, cls=c)
            assert np.abs(d[i] - ref).max() <= 1e-12 * np.abs(ref).max()
