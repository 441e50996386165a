"""Streamed serving of coarse-search requests (binding of the rc_server C
ABI) and the streamed-micrograph workload used to measure it.

Workload: Poisson arrivals of micrograph requests with 20-400 particles
each; box sizes drawn from {64, 128, 256} with coarse sizes {32, 64, 96};
1-4 classes per request; HEALPix order 2 x 24 psi (4608 orientations);
integer translations within radius 3-6 px.  Every (box, class) projector
volume is built once and stays device-resident.
"""
from __future__ import annotations

import ctypes as C
import math
import threading
import time

import numpy as np

from . import lib, synth

BOX_TO_COARSE = {64: 32, 128: 64, 256: 96}


class rc_server_config(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_size_t), ("pinned_budget_bytes", C.c_size_t),
                ("max_inflight", C.c_int), ("form", C.c_int)]


class rc_request(C.Structure):
    _fields_ = [("n_particles", C.c_int), ("n_classes", C.c_int), ("orientation_num", C.c_int),
                ("npix", C.c_int), ("full_npix", C.c_int),
                ("projectors", C.POINTER(lib.rc_projector_kernel)),
                ("trans_offset", C.POINTER(C.c_int)), ("trans_x", C.POINTER(C.c_float)),
                ("trans_y", C.POINTER(C.c_float)), ("eulers", C.POINTER(C.c_float)),
                ("real", C.POINTER(C.c_float)), ("imag", C.POINTER(C.c_float)),
                ("corr", C.POINTER(C.c_float)), ("coor_xy", C.POINTER(C.c_float)),
                ("keep_diff2", C.c_int)]


class rc_server_counters(C.Structure):
    _fields_ = [("completed", C.c_uint64), ("queued", C.c_int), ("inflight", C.c_int),
                ("peak_inflight", C.c_int), ("hbm_used", C.c_size_t), ("hbm_peak", C.c_size_t),
                ("hbm_budget", C.c_size_t), ("busy_s", C.c_double)]


def _bind():
    L = lib.load()
    if getattr(L, "_server_bound", False):
        return L
    vp = C.c_void_p
    L.rc_server_create.argtypes = [C.POINTER(rc_server_config), C.POINTER(vp)]
    L.rc_server_submit.argtypes = [vp, C.POINTER(rc_request), C.POINTER(C.c_uint64)]
    L.rc_server_wait.argtypes = [vp, C.c_uint64, C.POINTER(C.c_double),
                                 C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L.rc_server_diff2.argtypes = [vp, C.c_uint64, C.POINTER(C.c_float), C.c_size_t]
    L.rc_server_release.argtypes = [vp, C.c_uint64]
    L.rc_server_stats.argtypes = [vp, C.POINTER(rc_server_counters)]
    L.rc_server_destroy.argtypes = [vp]
    L._server_bound = True
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Request:
    """Host inputs of one request (kept alive until submitted)."""

    def __init__(self, projector_kernels, eulers, trans_offset, trans_x, trans_y, particles,
                 keep_diff2=False, box=None, classes=None):
        pb = particles
        self.pk = (lib.rc_projector_kernel * len(projector_kernels))(*projector_kernels)
        self.eulers = np.ascontiguousarray(eulers, np.float32).ravel()
        self.trans_offset = np.ascontiguousarray(trans_offset, np.int32)
        self.trans_x = np.ascontiguousarray(trans_x, np.float32)
        self.trans_y = np.ascontiguousarray(trans_y, np.float32)
        self.real = np.ascontiguousarray(pb.real, np.float32)
        self.imag = np.ascontiguousarray(pb.imag, np.float32)
        self.corr = np.ascontiguousarray(pb.corr, np.float32)
        self.coor = np.ascontiguousarray(pb.coor_xy, np.float32)
        self.P, self.C = pb.n, len(projector_kernels)
        self.O = self.eulers.size // 9
        self.npix, self.full = pb.npix, pb.full_npix
        self.keep = keep_diff2
        self.box, self.classes = box, classes

    def evals(self):
        return int(self.trans_offset[-1]) * self.C * self.O

    def flops(self):
        return 4.0 * self.npix * self.evals()

    def struct(self):
        r = rc_request()
        r.n_particles, r.n_classes, r.orientation_num = self.P, self.C, self.O
        r.npix, r.full_npix = self.npix, self.full
        r.projectors = self.pk
        r.trans_offset = self.trans_offset.ctypes.data_as(C.POINTER(C.c_int))
        r.trans_x, r.trans_y = _fp(self.trans_x), _fp(self.trans_y)
        r.eulers = _fp(self.eulers)
        r.real, r.imag, r.corr = _fp(self.real), _fp(self.imag), _fp(self.corr)
        r.coor_xy = _fp(self.coor)
        r.keep_diff2 = int(self.keep)
        return r


class Server:
    def __init__(self, form="bf16x3", hbm_budget=32 << 30, pinned_budget=2 << 30,
                 max_inflight=4):
        self.L = _bind()
        cfg = rc_server_config(int(hbm_budget), int(pinned_budget), int(max_inflight),
                               lib.FORMS[form] if isinstance(form, str) else int(form))
        self.h = C.c_void_p()
        lib.check(self.L.rc_server_create(C.byref(cfg), C.byref(self.h)))

    def submit(self, req: Request) -> int:
        t = C.c_uint64()
        s = req.struct()
        lib.check(self.L.rc_server_submit(self.h, C.byref(s), C.byref(t)))
        return t.value

    def wait(self, ticket, n_particles):
        best = np.zeros(n_particles, np.float64)
        idx = np.zeros(n_particles, np.int64)
        lat = C.c_double()
        lib.check(self.L.rc_server_wait(self.h, ticket,
                                        best.ctypes.data_as(C.POINTER(C.c_double)),
                                        idx.ctypes.data_as(C.POINTER(C.c_int64)),
                                        C.byref(lat)))
        return best, idx, lat.value

    def diff2(self, ticket, count):
        out = np.empty(count, np.float32)
        lib.check(self.L.rc_server_diff2(self.h, ticket, _fp(out), count))
        return out

    def release(self, ticket):
        lib.check(self.L.rc_server_release(self.h, ticket))

    def stats(self):
        c = rc_server_counters()
        lib.check(self.L.rc_server_stats(self.h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in c._fields_}

    def close(self):
        if self.h:
            lib.check(self.L.rc_server_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------------
# streamed workload
# --------------------------------------------------------------------------
class VolumeCache:
    """Device-resident projectors, one per (box, class id), built on first
    use and kept for the life of the cache."""

    def __init__(self, n_class_ids=4, seed=100):
        self.n, self.seed = n_class_ids, seed
        self.vols, self.proj = {}, {}

    def volume(self, box, cls):
        key = (box, cls)
        if key not in self.vols:
            r = BOX_TO_COARSE[box] // 2
            self.vols[key] = synth.fourier_volume_analytic(box, r, seed=self.seed + 31 * cls + box)
        return self.vols[key]

    def kernel(self, box, cls):
        key = (box, cls)
        if key not in self.proj:
            self.proj[key] = lib.Projector(self.volume(box, cls))
        y = BOX_TO_COARSE[box]
        return self.proj[key].kernel(y // 2 + 1, y, y // 2)

    def device_bytes(self):
        return sum(v.real.nbytes * 2 for v in self.vols.values())


def make_stream(n_requests, rate_hz, seed=7, particles=(20, 400), classes=(1, 4),
                radius=(3, 6), boxes=(64, 128, 256), order=2, n_psi=24, cache=None,
                snr=0.05):
    """Poisson arrival times and the host inputs of each request."""
    rng = np.random.default_rng(seed)
    cache = cache or VolumeCache(max(classes))
    eul = synth.orientation_set(order, n_psi)
    t = 0.0
    out = []
    for k in range(n_requests):
        t += rng.exponential(1.0 / rate_hz)
        box = int(rng.choice(boxes))
        y = BOX_TO_COARSE[box]
        g = synth.CoarseGeometry(imgY=y, maxR=y // 2)
        n_cls = int(rng.integers(classes[0], classes[1] + 1))
        cls_ids = list(range(n_cls))
        n_p = int(rng.integers(particles[0], particles[1] + 1))
        r = int(rng.integers(radius[0], radius[1] + 1))
        tx, ty, _ = synth.translation_set(r * r, box)
        vols = [cache.volume(box, c) for c in cls_ids]
        pb = synth.make_particles(vols, g, n_p, box, snr, seed=seed * 1000 + k,
                                  shift_range=min(2.0, r / 2))
        off = (np.arange(n_p + 1) * tx.size).astype(np.int32)
        req = Request([cache.kernel(box, c) for c in cls_ids], eul, off, np.tile(tx, n_p),
                      np.tile(ty, n_p), pb, box=box, classes=n_cls)
        out.append((t, req))
    return out, cache


def resample_stream(base, n_requests, rate_hz, seed=0):
    """n_requests Poisson arrivals at rate_hz, each request drawn uniformly
    from the generated stream `base` (its host inputs are submitted again in
    full, so the server does the whole work each time)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_hz, n_requests))
    pick = rng.integers(0, len(base), n_requests)
    return [(float(t[k]), base[int(pick[k])][1]) for k in range(n_requests)]


def run_stream(server, stream, time_scale=1.0):
    """Submit each request at its arrival time (from a submitter thread),
    wait for all, return per-request latencies (ms) and the wall time."""
    tickets = [None] * len(stream)
    err = []

    def submitter():
        t0 = time.perf_counter()
        for k, (ta, req) in enumerate(stream):
            dt = ta * time_scale - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            try:
                tickets[k] = server.submit(req)
            except Exception as e:  # surfaced below
                err.append(e)
                return

    t_start = time.perf_counter()
    th = threading.Thread(target=submitter)
    th.start()
    lat = []
    for k, (_, req) in enumerate(stream):
        while tickets[k] is None and not err:
            time.sleep(0.0005)
        if err:
            break
        _, _, ms = server.wait(tickets[k], req.P)
        lat.append(ms)
    th.join()
    if err:
        raise err[0]
    wall = time.perf_counter() - t_start
    return np.asarray(lat), wall


def summarize(lat_ms, stream, wall_s):
    evals = sum(r.evals() for _, r in stream)
    return dict(requests=len(stream), p50_ms=float(np.percentile(lat_ms, 50)),
                p99_ms=float(np.percentile(lat_ms, 99)), mean_ms=float(lat_ms.mean()),
                evals=evals, wall_s=wall_s, evals_per_s=evals / wall_s,
                offered_rate_hz=len(stream) / max(stream[-1][0], 1e-9))
