"""Python side of the CPU oracle.  TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liboracle.so`` (built from ``coarse_oracle.c`` by the Makefile
beside it) and wraps it for tests, the smoke check and the CPU-baseline leg
of bench.py.  The product library never imports this module.

Also holds the error metrics the reference reports (check.cuh:84-148):
absolute error on every element, relative error only where the reference
value exceeds 1e-12 in magnitude, mean and max of each.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-B", "-C", _HERE], check=True, capture_output=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.oracle_project_f64.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_float, C.c_int, _f32p, C.c_int, _f64p, _f64p,
                                         _f64p, _f64p]
        L.oracle_project_f64.restype = None
        L.oracle_project_many_f64.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int,
                                               C.c_int, C.c_float, C.c_int, _f32p, C.c_int,
                                               C.c_int, _f64p, _f64p, _f64p, _f64p]
        L.oracle_project_many_f64.restype = None
        L.oracle_translate_f64.argtypes = [C.c_int, _f64p, _f64p, C.c_float, C.c_float,
                                           _f32p, _f32p, _f64p, _f64p]
        L.oracle_translate_f64.restype = None
        L.oracle_diff2_f64.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_float, C.c_int, _f32p, C.c_int, _f32p, _f32p, C.c_int,
                                       _f32p, _f32p, _f32p, _f64p, _f64p, C.c_int, C.c_int,
                                       C.c_void_p, _f64p]
        L.oracle_diff2_f64.restype = None
        L.oracle_diff2_f32.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_float, C.c_int, _f32p, C.c_int, _f32p, _f32p, C.c_int,
                                       _f32p, _f32p, _f32p, _f32p, C.c_int, C.c_int, _f32p, C.c_int]
        L.oracle_diff2_f32.restype = C.c_int
        L.oracle_cc_f64.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, _f32p, C.c_int,
                                    _f32p, _f32p, C.c_int, _f32p, _f32p, _f32p, C.c_int, _f64p]
        L.oracle_cc_f64.restype = None
        _i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        L.oracle_fine_f64.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, _f32p,
                                      _f32p, _f32p, _f32p, _f32p, _f32p, C.c_int, C.c_double,
                                      C.c_int, _i64p, _i64p, _i64p, _i64p, _f64p]
        L.oracle_fine_f64.restype = None
        _LIB = L
    return _LIB


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _vol_args(v):
    return (_c32(v.real), _c32(v.imag), v.mdlX, v.mdlY, v.mdlZ, v.initY, v.initZ, float(v.pad))


def project(v, maxR2_padded, e9, xs, ys):
    """float64 projection of one orientation at pixels (xs, ys) -> complex128."""
    xs, ys = _c64(xs), _c64(ys)
    re = np.zeros(xs.size)
    im = np.zeros(xs.size)
    lib().oracle_project_f64(*_vol_args(v), int(maxR2_padded), _c32(e9), xs.size, xs, ys, re, im)
    return re + 1j * im


def translate(xs, ys, tx, ty, re, im):
    xs, ys = _c64(xs), _c64(ys)
    o_re = np.zeros(xs.size)
    o_im = np.zeros(xs.size)
    lib().oracle_translate_f64(xs.size, xs, ys, float(tx), float(ty), _c32(re), _c32(im), o_re, o_im)
    return o_re + 1j * o_im


def diff2_f64(v, maxR2_padded, eulers, trans_x, trans_y, re, im, corr, xs, ys,
              n_proj=None, init=None):
    """float64 coarse diff2 of one image: returns ``[O, T]``."""
    eulers = _c32(eulers).reshape(-1, 9)
    O, T = eulers.shape[0], int(np.size(trans_x))
    xs, ys = _c64(xs), _c64(ys)
    n = xs.size
    out = np.zeros(O * T)
    init_p = None
    if init is not None:
        init = _c64(init).ravel()
        init_p = init.ctypes.data
    lib().oracle_diff2_f64(*_vol_args(v), int(maxR2_padded), eulers.ravel(), O, _c32(trans_x),
                           _c32(trans_y), T, _c32(re), _c32(im), _c32(corr), xs, ys, n,
                           n if n_proj is None else int(n_proj), init_p, out)
    return out.reshape(O, T)


def diff2_f32(v, maxR2_padded, eulers, trans_x, trans_y, re, im, corr, coor_xy, n_proj,
              out=None, threads=0):
    """float32 port of the CUDA-core loop; adds into ``out`` ``[O*T]``.
    Returns (out, threads used)."""
    eulers = _c32(eulers).reshape(-1, 9)
    O, T = eulers.shape[0], int(np.size(trans_x))
    coor = _c32(coor_xy).ravel()
    n = coor.size // 2
    if out is None:
        out = np.zeros(O * T, np.float32)
    used = lib().oracle_diff2_f32(*_vol_args(v), int(maxR2_padded), eulers.ravel(), O,
                                  _c32(trans_x), _c32(trans_y), T, _c32(re), _c32(im),
                                  _c32(corr), coor, n, int(n_proj), out, int(threads))
    return out, used


def cc_f64(v, maxR2_padded, imgX, imgY, maxR, eulers, trans_x, trans_y, re, im, corr):
    """float64 cross-correlation score of one unfiltered image: ``[O, T]``."""
    eulers = _c32(eulers).reshape(-1, 9)
    O, T = eulers.shape[0], int(np.size(trans_x))
    n = int(np.size(re))
    out = np.zeros(O * T)
    lib().oracle_cc_f64(*_vol_args(v), int(maxR2_padded), imgX, imgY, maxR, eulers.ravel(), O,
                        _c32(trans_x), _c32(trans_y), T, _c32(re), _c32(im), _c32(corr), n, out)
    return out.reshape(O, T)


def fine_f64(v, maxR2_padded, imgX, imgY, maxR, eulers, trans_x, trans_y, re, im, corr,
             rot_idx, trans_idx, job_idx, job_num, sum_init=0.0, n_out=None):
    """float64 fine-pass diff2 of one unfiltered image over a job list;
    returns ``out`` with out[job_idx[b] + j] filled (NaN elsewhere)."""
    i64 = lambda a: np.ascontiguousarray(a, np.int64)
    rot_idx, trans_idx, job_idx, job_num = map(i64, (rot_idx, trans_idx, job_idx, job_num))
    n = int(np.size(re))
    if n_out is None:
        n_out = int((job_idx + job_num).max()) if job_num.size else 0
    out = np.full(n_out, np.nan)
    lib().oracle_fine_f64(*_vol_args(v), int(maxR2_padded), imgX, imgY, maxR,
                          _c32(eulers).ravel(), _c32(trans_x), _c32(trans_y), _c32(re), _c32(im),
                          _c32(corr), n, float(sum_init), int(job_num.size), rot_idx, trans_idx,
                          job_idx, job_num, out)
    return out


def problem_diff2(prob, i, cls=0, orient=None, init=None):
    """float64 diff2 ``[O, T_i]`` of particle ``i`` against class ``cls`` of a
    synthetic problem (filtered pixel order, tail = out-of-radius)."""
    pb = prob.particles
    tx, ty = prob.trans_of(i)
    eul = prob.eulers if orient is None else prob.eulers[orient]
    return diff2_f64(prob.vols[cls], prob.geom.maxR2_padded, eul, tx, ty, pb.real[i], pb.imag[i],
                     pb.corr[i], pb.coor_xy[:, 0], pb.coor_xy[:, 1], n_proj=pb.npix, init=init)


def projections(v, maxR2_padded, eulers, xs, ys):
    """float64 projections ``[O, n]`` (complex128) of every orientation at
    pixels (xs, ys): the projection step of ``diff2_f64``."""
    eulers = _c32(eulers).reshape(-1, 9)
    O = eulers.shape[0]
    xs, ys = _c64(xs), _c64(ys)
    re = np.zeros(O * xs.size)
    im = np.zeros(O * xs.size)
    lib().oracle_project_many_f64(*_vol_args(v), int(maxR2_padded), eulers.ravel(), O, xs.size,
                                  xs, ys, re, im)
    return (re + 1j * im).reshape(O, xs.size)


def diff2_expanded(P, trans_x, trans_y, re, im, corr, xs, ys, n_proj):
    """float64 coarse diff2 ``[O, T]`` of one image in the expanded form
    sum w/2 |P|^2 + sum w/2 |X|^2 - sum w Re(P conj(X e^{i phi})), with the
    projections ``P [O, n_proj]`` from ``projections`` and the pixel sums
    as float64 matrix products.  The same quantity as ``diff2_f64`` (the
    direct form sum w/2 |P - Xs|^2), which pins it in the CPU tests; it is
    here because it costs one complex GEMM per image instead of a sincos
    per (orientation, translation, pixel)."""
    xs, ys = _c64(xs), _c64(ys)
    w = np.asarray(corr, np.float64)
    X = np.asarray(re, np.float64) + 1j * np.asarray(im, np.float64)
    tx = np.asarray(trans_x, np.float32).astype(np.float64)
    ty = np.asarray(trans_y, np.float32).astype(np.float64)
    xn = 0.5 * float(np.sum(w * (X.real ** 2 + X.imag ** 2)))
    k = int(n_proj)
    phi = tx[:, None] * xs[None, :k] + ty[:, None] * ys[None, :k]
    A = (w[:k] * X[:k])[None, :] * np.exp(1j * phi)          # [T, k]: w Xs
    pn = (P.real ** 2 + P.imag ** 2) @ (0.5 * w[:k])          # [O]
    cross = (P @ np.conj(A).T).real                           # [O, T]
    return pn[:, None] + xn - cross


def problem_diff2_expanded(prob, particles, cls=0, P=None):
    """``diff2_expanded`` of the given particles of a synthetic problem
    against class ``cls``: {i: [O, T_i]}.  ``P`` may hold the class's
    projections from an earlier call (``projections`` at the in-radius
    pixels)."""
    pb = prob.particles
    xs, ys = pb.coor_xy[:, 0], pb.coor_xy[:, 1]
    if P is None:
        P = projections(prob.vols[cls], prob.geom.maxR2_padded, prob.eulers, xs[:pb.npix],
                        ys[:pb.npix])
    out = {}
    for i in particles:
        tx, ty = prob.trans_of(i)
        out[int(i)] = diff2_expanded(P, tx, ty, pb.real[i], pb.imag[i], pb.corr[i], xs, ys,
                                     pb.npix)
    return out


def error_stats(a, ref):
    """Reference error metrics (check.cuh:96-111)."""
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(ref, np.float64).ravel()
    ab = np.abs(a - b)
    m = np.abs(b) > 1e-12
    rel = ab[m] / np.abs(b[m])
    return dict(mean_abs=float(ab.mean()) if ab.size else 0.0,
                max_abs=float(ab.max()) if ab.size else 0.0,
                mean_rel=float(rel.mean()) if rel.size else 0.0,
                max_rel=float(rel.max()) if rel.size else 0.0,
                count=int(ab.size), count_rel=int(rel.size))


def argmin_flat(d):
    """Lowest flat index of the minimum (ties broken by the lower index)."""
    d = np.asarray(d).ravel()
    return int(np.argmin(d))


# ---------------------------------------------------------------------------
# coarse weights and significance (the step after the search)
# ---------------------------------------------------------------------------
def significance_threshold(w, fraction, total=None):
    """RELION's significance rule on one particle's weights: sort ascending,
    inclusive cumulative sums c, threshold (1 - fraction) * total, index
    i + 1 of the pair c[i] <= threshold < c[i + 1] (0 if none), and the
    weight at that index (find_threshold_idx_in_cumulative,
    cuda_device_utils.cuh:251-262; array_over_threshold :238-249).  Written
    as the literal scan over pairs so it restates the reference's rule.
    Returns (significant weight, number of weights >= it)."""
    s = np.sort(np.asarray(w, np.float64).ravel())
    c = np.cumsum(s)
    total = c[-1] if total is None else float(total)
    thr = (1.0 - fraction) * total
    idx = 0
    hit = np.nonzero((c[:-1] <= thr) & (thr < c[1:]))[0]
    if hit.size:
        idx = int(hit[0]) + 1
    ws = s[idx]
    return ws, int(np.count_nonzero(np.asarray(w).ravel() >= ws))


def coarse_weights(diff2, min_diff2, n_classes, orientation_num, T, fraction,
                   log_prior_orient=None, log_prior_offset=None):
    """float64 weights of one particle from its coarse diff2 ``[C*O*T]``
    (utilities.h:405-459 semantics): lw = lp_orient + lp_offset + min - diff2,
    w = exp(lw - max lw); samples with a -inf log prior or diff2 below
    (float)min get w = 0.  Returns (w, sum, significant weight, count,
    max lw)."""
    C, O = n_classes, orientation_num
    d = np.asarray(diff2, np.float64).reshape(C, O, T)
    lpo = (np.zeros((C, O, 1)) if log_prior_orient is None
           else np.asarray(log_prior_orient, np.float64).reshape(C, O, 1))
    lpt = (np.zeros((C, 1, T)) if log_prior_offset is None
           else np.asarray(log_prior_offset, np.float64).reshape(C, 1, T))
    ok = (d >= np.float32(min_diff2)) & np.isfinite(lpo) & np.isfinite(lpt)
    with np.errstate(invalid="ignore"):
        lw = lpo + lpt + (float(min_diff2) - d)
    if not ok.any():
        w = np.zeros_like(d)
        ws, n = significance_threshold(w, fraction, 0.0)
        return w.ravel(), 0.0, ws, n, -np.inf
    lmax = lw[ok].max()
    w = np.where(ok, np.exp(np.where(ok, lw - lmax, 0.0)), 0.0)
    total = float(w.sum())
    ws, n = significance_threshold(w, fraction, total)
    return w.ravel(), total, ws, n, lmax
