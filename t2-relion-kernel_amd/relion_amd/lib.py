"""ctypes binding of ``librelion_coarse_amd.so`` (the C ABI declared in
``include/relion_coarse.h``).

Device buffers are torch tensors on ``cuda`` (PyTorch is used for device
memory and streams only); every compute call goes through the C ABI into
the gfx950 HIP kernels.  If the shared library is missing the import of
this module raises: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "librelion_coarse_amd.so")

FORM_FP32, FORM_BF16X3, FORM_FP16, FORM_FP32_DIRECT = 0, 1, 2, 3
FORMS = {"fp32": FORM_FP32, "bf16x3": FORM_BF16X3, "fp16": FORM_FP16,
         "fp32_direct": FORM_FP32_DIRECT}

RC_OK = 0
STATUS = {0: "RC_OK", 1: "RC_ERR_INVALID", 2: "RC_ERR_UNSUPPORTED", 3: "RC_ERR_NO_MEMORY",
          4: "RC_ERR_DEVICE", 5: "RC_ERR_LIMIT", 6: "RC_ERR_IO"}

# every symbol the header declares
EXPORTS = [
    "rc_projector_create", "rc_projector_create_from_file", "rc_projector_destroy",
    "rc_projector_dims", "rc_projector_make_kernel", "rc_run_diff2_kernel_coarse",
    "rc_coarse_workspace_bytes", "rc_coarse_run", "rc_plan_create", "rc_plan_execute",
    "rc_plan_destroy", "rc_plan_device_bytes", "rc_plan_tile", "rc_set_variant_override",
    "rc_variant_count", "rc_variant_info", "rc_last_error", "rc_version", "rc_plan_timing",
    "rc_plan_timing_read", "rc_autotune", "rc_plan_capture", "rc_plan_replay",
    "rc_server_create", "rc_server_submit", "rc_server_wait", "rc_server_diff2",
    "rc_server_release", "rc_server_stats", "rc_server_destroy", "rc_coarse_data_load",
    "rc_coarse_data_free", "rc_variant_name", "rc_coarse_weights",
    "rc_run_diff2_kernel_fine", "rc_scratch_release", "rc_coarse_split",
    "rc_filter_coarse_order", "rc_filter_coarse_images", "rc_run_diff2_kernel_coarse_noalloc",
]


class RcError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class rc_projector_kernel(C.Structure):
    _fields_ = [("proj", C.c_void_p), ("imgX", C.c_int), ("imgY", C.c_int), ("imgZ", C.c_int),
                ("maxR", C.c_int), ("maxR2_padded", C.c_int), ("padding_factor", C.c_float)]


class rc_batch_desc(C.Structure):
    _fields_ = [("n_particles", C.c_int), ("n_classes", C.c_int), ("orientation_num", C.c_int),
                ("npix", C.c_int), ("full_npix", C.c_int), ("form", C.c_int),
                ("projectors", C.POINTER(rc_projector_kernel)),
                ("host_trans_offset", C.POINTER(C.c_int)),
                ("d_trans_offset", C.c_void_p), ("d_trans_x", C.c_void_p),
                ("d_trans_y", C.c_void_p), ("d_eulers", C.c_void_p), ("d_real", C.c_void_p),
                ("d_imag", C.c_void_p), ("d_corr", C.c_void_p), ("d_coor_xy", C.c_void_p),
                ("d_diff2_init", C.c_void_p), ("d_diff2", C.c_void_p),
                ("d_best_diff2", C.c_void_p), ("d_best_index", C.c_void_p),
                ("do_cc", C.c_int)]


class rc_weights_desc(C.Structure):
    _fields_ = [("n_particles", C.c_int), ("n_classes", C.c_int), ("orientation_num", C.c_int),
                ("host_trans_offset", C.POINTER(C.c_int)), ("d_trans_offset", C.c_void_p),
                ("d_diff2", C.c_void_p), ("d_min_diff2", C.c_void_p),
                ("d_log_prior_orient", C.c_void_p), ("log_prior_orient_stride", C.c_longlong),
                ("d_log_prior_offset", C.c_void_p), ("adaptive_fraction", C.c_double),
                ("d_weights", C.c_void_p), ("d_sum_weight", C.c_void_p),
                ("d_significant_weight", C.c_void_p), ("d_n_significant", C.c_void_p),
                ("d_significant", C.c_void_p), ("d_max_log_weight", C.c_void_p)]


_lib = None


def load():
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C {PKG_ROOT}` "
                          "(or __graft_entry__.build()); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    vp, i, f, sz = C.c_void_p, C.c_int, C.c_float, C.c_size_t
    fp = C.POINTER(C.c_float)
    L.rc_projector_create.argtypes = [i, i, i, i, i, i, f, fp, fp, C.POINTER(vp)]
    L.rc_projector_create_from_file.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.rc_projector_destroy.argtypes = [vp]
    L.rc_projector_dims.argtypes = [vp] + [C.POINTER(i)] * 6 + [C.POINTER(f)]
    L.rc_projector_make_kernel.argtypes = [vp, i, i, i, i, C.POINTER(rc_projector_kernel)]
    L.rc_run_diff2_kernel_coarse.argtypes = [C.POINTER(rc_projector_kernel), vp, vp, vp, vp, vp,
                                             vp, vp, vp, f, C.c_ulong, C.c_ulong, C.c_ulong, vp,
                                             i, i]
    L.rc_coarse_workspace_bytes.argtypes = [i, i, i, i, i]
    L.rc_coarse_workspace_bytes.restype = sz
    L.rc_run_diff2_kernel_coarse_noalloc.argtypes = L.rc_run_diff2_kernel_coarse.argtypes
    L.rc_scratch_release.argtypes = [vp]
    L.rc_coarse_split.argtypes = [i, i, i, i, C.POINTER(i), C.POINTER(i)]
    L.rc_filter_coarse_order.argtypes = [C.POINTER(rc_projector_kernel), i, C.c_void_p,
                                         C.POINTER(i)]
    L.rc_filter_coarse_images.argtypes = [C.POINTER(rc_projector_kernel), i, i] + [vp] * 8
    L.rc_coarse_run.argtypes = [i, i, i, i, i, vp, vp, vp, vp, vp,
                                C.POINTER(rc_projector_kernel), vp, vp, vp, vp, vp, vp, sz]
    L.rc_plan_create.argtypes = [C.POINTER(rc_batch_desc), C.POINTER(vp)]
    L.rc_coarse_weights.argtypes = [C.POINTER(rc_weights_desc), vp]
    ul = C.c_ulong
    L.rc_run_diff2_kernel_fine.argtypes = [C.POINTER(rc_projector_kernel)] + [vp] * 8 + [
        ul, f, ul, ul, ul] + [vp] * 4 + [i, vp]
    L.rc_plan_execute.argtypes = [vp, C.POINTER(rc_batch_desc), vp]
    L.rc_plan_destroy.argtypes = [vp]
    L.rc_plan_device_bytes.argtypes = [vp]
    L.rc_plan_device_bytes.restype = sz
    L.rc_plan_tile.argtypes = [vp, C.POINTER(i), C.POINTER(i), C.POINTER(i)]
    L.rc_autotune.argtypes = [C.POINTER(rc_batch_desc), vp, i, C.POINTER(i), fp]
    L.rc_plan_timing.argtypes = [vp, i]
    L.rc_plan_capture.argtypes = [vp, C.POINTER(rc_batch_desc), vp]
    L.rc_plan_replay.argtypes = [vp, vp]
    L.rc_plan_timing_read.argtypes = [vp, fp, i, C.POINTER(i)]
    L.rc_set_variant_override.argtypes = [i]
    L.rc_variant_count.argtypes = [i]
    L.rc_variant_info.argtypes = [i, i, C.POINTER(i), C.POINTER(i), C.POINTER(i)]
    L.rc_variant_name.argtypes = [i, i]
    L.rc_variant_name.restype = C.c_char_p
    L.rc_last_error.restype = C.c_char_p
    L.rc_version.restype = C.c_char_p
    _lib = L
    return L


def check(status):
    if status != RC_OK:
        raise RcError(status, load().rc_last_error().decode())


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _f32ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Projector:
    """Device-resident padded Fourier volume (owns its device memory)."""

    def __init__(self, vol=None, path=None):
        L = load()
        h = C.c_void_p()
        if path is not None:
            check(L.rc_projector_create_from_file(path.encode(), C.byref(h)))
        else:
            re = np.ascontiguousarray(vol.real, np.float32)
            im = np.ascontiguousarray(vol.imag, np.float32)
            check(L.rc_projector_create(vol.mdlX, vol.mdlY, vol.mdlZ, vol.initY, vol.initZ,
                                        vol.maxR, float(vol.pad), _f32ptr(re), _f32ptr(im),
                                        C.byref(h)))
        self.handle = h

    def dims(self):
        v = [C.c_int() for _ in range(6)]
        pad = C.c_float()
        check(load().rc_projector_dims(self.handle, *[C.byref(x) for x in v], C.byref(pad)))
        return tuple(x.value for x in v) + (pad.value,)

    def kernel(self, imgX, imgY, imgMaxR, imgZ=1):
        k = rc_projector_kernel()
        check(load().rc_projector_make_kernel(self.handle, imgX, imgY, imgZ, imgMaxR, C.byref(k)))
        return k

    def close(self):
        if self.handle:
            check(load().rc_projector_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_diff2_kernel_coarse(pk, trans_x, trans_y, corr, real, imag, eulers, diff2s,
                            orientation_num, translation_num, image_size, stream=None,
                            do_CC=False, data_is_3D=False):
    """Mirror of runDiff2KernelCoarse: ``diff2s[o*T+t] +=`` score (torch tensors)."""
    check(load().rc_run_diff2_kernel_coarse(
        C.byref(pk), _ptr(trans_x), _ptr(trans_y), None, _ptr(corr), _ptr(real), _ptr(imag),
        _ptr(eulers), _ptr(diff2s), 1.0, orientation_num, translation_num, image_size,
        _stream(stream), int(do_CC), int(data_is_3D)))


def run_diff2_kernel_coarse_noalloc(pk, trans_x, trans_y, corr, real, imag, eulers, diff2s,
                                    orientation_num, translation_num, image_size, stream=None,
                                    do_CC=False, data_is_3D=False):
    """rc_run_diff2_kernel_coarse_noalloc: the same call with the library's
    static scratch (no allocation; overlapping calls on one GPU race)."""
    check(load().rc_run_diff2_kernel_coarse_noalloc(
        C.byref(pk), _ptr(trans_x), _ptr(trans_y), None, _ptr(corr), _ptr(real), _ptr(imag),
        _ptr(eulers), _ptr(diff2s), 1.0, orientation_num, translation_num, image_size,
        _stream(stream), int(do_CC), int(data_is_3D)))


class CoarseKernel:
    """Mirror of the reference's coarse kernel structs: construct with the
    problem shape, query the workspace, then ``run``."""

    def __init__(self, translation_num, orientation_num, image_size, full_image_size=0,
                 form=FORM_FP32):
        self.T, self.O, self.npix = translation_num, orientation_num, image_size
        self.full = full_image_size
        self.form = FORMS[form] if isinstance(form, str) else form

    def get_workspace_size_bytes(self):
        return int(load().rc_coarse_workspace_bytes(self.T, self.O, self.npix, self.full,
                                                    self.form))

    def run(self, eulers, trans_x, trans_y, real, imag, pk, corr, diff2s, diff2s_dest,
            stream=None, coor_xy=None, workspace=None):
        import torch
        if workspace is None:
            workspace = torch.empty(max(self.get_workspace_size_bytes(), 1), dtype=torch.uint8,
                                    device=real.device)
        check(load().rc_coarse_run(self.T, self.O, self.npix, self.full, self.form, _ptr(eulers),
                                   _ptr(trans_x), _ptr(trans_y), _ptr(real), _ptr(imag),
                                   C.byref(pk), _ptr(corr), _ptr(diff2s), _ptr(diff2s_dest),
                                   _stream(stream), _ptr(coor_xy), _ptr(workspace),
                                   workspace.numel()))


def coarse_split(translation_num, orientation_num, image_size, form="fp32"):
    """(slices, 32-pixel chunks per slice) of the single-image calls' split-K."""
    f = FORMS[form] if isinstance(form, str) else form
    k, c = C.c_int(), C.c_int()
    check(load().rc_coarse_split(translation_num, orientation_num, image_size, f, C.byref(k),
                                 C.byref(c)))
    return k.value, c.value


def scratch_release(stream=None):
    check(load().rc_scratch_release(_stream(stream)))


def geometry_view(imgX, imgY, maxR, pad, maxR2_padded=None):
    """A projector-kernel view holding only image geometry (no volume): enough
    for the in-radius partition calls."""
    k = rc_projector_kernel()
    k.proj = None
    k.imgX, k.imgY, k.imgZ, k.maxR = imgX, imgY, 1, maxR
    k.padding_factor = pad
    k.maxR2_padded = int(maxR * maxR * pad * pad) if maxR2_padded is None else maxR2_padded
    return k


def filter_order(view, image_size):
    """rc_filter_coarse_order: (filtered position -> source pixel, inside count)."""
    order = np.empty(image_size, np.int32)
    n = C.c_int()
    check(load().rc_filter_coarse_order(C.byref(view), int(image_size),
                                        order.ctypes.data_as(C.c_void_p), C.byref(n)))
    return order, n.value


def filter_images(view, image_size, fimg, corr, stream=None):
    """rc_filter_coarse_images on torch device tensors: fimg [n][2*image_size]
    (real block, imaginary block), corr [n][image_size].  Returns real, imag,
    corr [n][image_size], coor_xy [image_size*2], filtered size [n]."""
    import torch
    n = fimg.numel() // (2 * image_size)
    dev = fimg.device
    out = dict(real=torch.empty(n, image_size, dtype=torch.float32, device=dev),
               imag=torch.empty(n, image_size, dtype=torch.float32, device=dev),
               corr=torch.empty(n, image_size, dtype=torch.float32, device=dev),
               coor_xy=torch.empty(2 * image_size, dtype=torch.float32, device=dev),
               filtered_size=torch.empty(n, dtype=torch.int32, device=dev))
    check(load().rc_filter_coarse_images(
        C.byref(view), int(image_size), int(n), _ptr(fimg), _ptr(corr), _ptr(out["real"]),
        _ptr(out["imag"]), _ptr(out["corr"]), _ptr(out["coor_xy"]), _ptr(out["filtered_size"]),
        _stream(stream)))
    return out


def variants(form):
    L = load()
    f = FORMS[form] if isinstance(form, str) else form
    out = []
    for v in range(L.rc_variant_count(f)):
        tm, tn, fu = C.c_int(), C.c_int(), C.c_int()
        check(L.rc_variant_info(f, v, C.byref(tm), C.byref(tn), C.byref(fu)))
        out.append((v, tm.value, tn.value))
    return out


def variant_name(form, v):
    """Kernel template and shape of variant ``v`` (as a profiler names it)."""
    f = FORMS[form] if isinstance(form, str) else form
    name = load().rc_variant_name(f, int(v))
    if name is None:
        raise RcError(f"no variant {v} for form {form}")
    return name.decode()


def set_variant_override(v):
    check(load().rc_set_variant_override(int(v)))


class BatchPlan:
    """A batched coarse search: particles x classes x orientations x ragged
    translations, one plan per shape."""

    def __init__(self, projector_kernels, trans_offset, n_particles, orientation_num, npix,
                 full_npix, form, cc=False):
        self.L = load()
        self.form = FORMS[form] if isinstance(form, str) else form
        self.pk = (rc_projector_kernel * len(projector_kernels))(*projector_kernels)
        self.trans_offset = np.ascontiguousarray(trans_offset, np.int32)
        self.P, self.C, self.O = n_particles, len(projector_kernels), orientation_num
        self.npix, self.full = npix, full_npix
        self.cc = bool(cc)
        self.handle = C.c_void_p()
        d = self._desc(dict())
        check(self.L.rc_plan_create(C.byref(d), C.byref(self.handle)))

    def _desc(self, t):
        d = rc_batch_desc()
        d.n_particles, d.n_classes, d.orientation_num = self.P, self.C, self.O
        d.npix, d.full_npix, d.form = self.npix, self.full, self.form
        d.projectors = self.pk
        d.host_trans_offset = self.trans_offset.ctypes.data_as(C.POINTER(C.c_int))
        d.do_cc = int(self.cc)
        for k in ("trans_offset", "trans_x", "trans_y", "eulers", "real", "imag", "corr",
                  "coor_xy", "diff2_init", "diff2", "best_diff2", "best_index"):
            v = t.get(k)
            setattr(d, "d_" + k, None if v is None else v.data_ptr())
        return d

    def tile(self):
        v, tm, tn = C.c_int(), C.c_int(), C.c_int()
        check(self.L.rc_plan_tile(self.handle, C.byref(v), C.byref(tm), C.byref(tn)))
        return v.value, tm.value, tn.value

    def device_bytes(self):
        return int(self.L.rc_plan_device_bytes(self.handle))

    def timing(self, n_slots):
        """Record HIP events around each cross-term launch (n_slots ring)."""
        check(self.L.rc_plan_timing(self.handle, int(n_slots)))

    def timing_read(self, n):
        ms = (C.c_float * n)()
        k = C.c_int()
        check(self.L.rc_plan_timing_read(self.handle, ms, n, C.byref(k)))
        return list(ms[:k.value])

    def autotune(self, tensors, iters=3, stream=None):
        """Time every variant on this shape; later plans of the shape use the
        fastest.  Returns (variant, ms of the cross-term launch)."""
        d = self._desc(tensors)
        v, ms = C.c_int(), C.c_float()
        check(self.L.rc_autotune(C.byref(d), _stream(stream), int(iters), C.byref(v),
                                 C.byref(ms)))
        return v.value, ms.value

    def execute(self, tensors, stream=None):
        d = self._desc(tensors)
        self._last_desc = d
        check(self.L.rc_plan_execute(self.handle, C.byref(d), _stream(stream)))

    def capture(self, tensors, stream):
        """Capture one execute into a hipGraph held by the plan."""
        d = self._desc(tensors)
        self._graph_desc = d
        check(self.L.rc_plan_capture(self.handle, C.byref(d), _stream(stream)))

    def replay(self, stream=None):
        check(self.L.rc_plan_replay(self.handle, _stream(stream)))

    def close(self):
        if self.handle:
            check(self.L.rc_plan_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def output_offsets(trans_offset, n_classes, orientation_num):
    """Start of each particle's block in the batched diff2 output."""
    return np.asarray(trans_offset, np.int64) * n_classes * orientation_num


def problem_tensors(prob, device="cuda"):
    """Upload the inputs of a synthetic problem (relion_amd.synth.CoarseProblem)."""
    import torch
    pb = prob.particles
    t = lambda a, dt=torch.float32: torch.as_tensor(np.ascontiguousarray(a)).to(
        device=device, dtype=dt)
    total = int(prob.trans_offset[-1]) * prob.n_classes * prob.orientation_num
    return dict(
        trans_offset=t(prob.trans_offset, torch.int32), trans_x=t(prob.trans_x),
        trans_y=t(prob.trans_y), eulers=t(prob.eulers.ravel()), real=t(pb.real.ravel()),
        imag=t(pb.imag.ravel()), corr=t(pb.corr.ravel()), coor_xy=t(pb.coor_xy.ravel()),
        diff2=torch.empty(total, dtype=torch.float32, device=device),
        best_diff2=torch.empty(prob.n_particles, dtype=torch.float64, device=device),
        best_index=torch.empty(prob.n_particles, dtype=torch.int64, device=device))


def run_diff2_kernel_fine(pk, eulers, real, imag, trans_x, trans_y, corr, diff2, image_size,
                          sum_init, orientation_num, translation_num, rot_idx, trans_idx,
                          job_idx, job_num, data_is_3D=False, stream=None):
    """rc_run_diff2_kernel_fine on torch device tensors (index arrays int64,
    passed as unsigned long)."""
    import torch
    L = load()
    p = lambda x: None if x is None else x.data_ptr()
    s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(L.rc_run_diff2_kernel_fine(C.byref(pk), p(eulers), p(real), p(imag), p(trans_x),
                                     p(trans_y), None, p(corr), p(diff2), int(image_size),
                                     float(sum_init), int(orientation_num), int(translation_num),
                                     int(job_num.numel()), p(rot_idx), p(trans_idx), p(job_idx),
                                     p(job_num), 1 if data_is_3D else 0, s))


def coarse_weights(trans_offset, d_trans_offset, n_classes, orientation_num, diff2, min_diff2,
                   log_prior_orient=None, orient_stride=0, log_prior_offset=None,
                   fraction=0.999, weights=None, mask=True, stream=None):
    """rc_coarse_weights on torch device tensors; returns a dict with
    ``weights`` (float32, diff2 layout), ``sum`` (float64 [P]),
    ``significant_weight`` (float32 [P]), ``n_significant`` (int32 [P]),
    ``mask`` (uint8, diff2 layout, or None) and ``max_log_weight``.  ``mask``
    may be False (no mask), True (allocated here) or a uint8 device tensor of
    at least ``diff2.numel()`` bytes."""
    import torch
    L = load()
    P = len(trans_offset) - 1
    dev = diff2.device
    if weights is None:
        weights = torch.empty_like(diff2)
    out = dict(weights=weights,
               sum=torch.empty(P, dtype=torch.float64, device=dev),
               significant_weight=torch.empty(P, dtype=torch.float32, device=dev),
               n_significant=torch.empty(P, dtype=torch.int32, device=dev),
               mask=(mask if isinstance(mask, torch.Tensor) else
                     torch.empty(diff2.numel(), dtype=torch.uint8, device=dev) if mask else None),
               max_log_weight=torch.empty(P, dtype=torch.float64, device=dev))
    off = np.ascontiguousarray(trans_offset, np.int32)
    d = rc_weights_desc()
    d.n_particles, d.n_classes, d.orientation_num = P, n_classes, orientation_num
    d.host_trans_offset = off.ctypes.data_as(C.POINTER(C.c_int))
    d.d_trans_offset = d_trans_offset.data_ptr()
    d.d_diff2 = diff2.data_ptr()
    d.d_min_diff2 = min_diff2.data_ptr()
    d.d_log_prior_orient = log_prior_orient.data_ptr() if log_prior_orient is not None else None
    d.log_prior_orient_stride = int(orient_stride)
    d.d_log_prior_offset = log_prior_offset.data_ptr() if log_prior_offset is not None else None
    d.adaptive_fraction = float(fraction)
    d.d_weights = weights.data_ptr()
    d.d_sum_weight = out["sum"].data_ptr()
    d.d_significant_weight = out["significant_weight"].data_ptr()
    d.d_n_significant = out["n_significant"].data_ptr()
    d.d_significant = out["mask"].data_ptr() if out["mask"] is not None else None
    d.d_max_log_weight = out["max_log_weight"].data_ptr()
    s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(L.rc_coarse_weights(C.byref(d), s))
    return out


def plan_for_problem(prob, projectors, form, cc=False):
    g = prob.geom
    pks = [p.kernel(g.imgX, g.imgY, g.maxR) for p in projectors]
    pb = prob.particles
    return BatchPlan(pks, prob.trans_offset, prob.n_particles, prob.orientation_num, pb.npix,
                     pb.full_npix, form, cc=cc)
