"""Replay of the reference's on-disk inputs.

Reads the projector dumps (``projector_proc_1_<iter>.dat``, read natively by
``rc_projector_create_from_file``) and the coarse-particle dumps
(``<iter>/coarse_data<d>00_1.dat``, read natively by ``rc_coarse_data_load``)
and runs them through ``rc_run_diff2_kernel_coarse`` the way the reference
harness drives ``runDiff2KernelCoarse`` (main.cu:226-331).  The writers
produce files in the same formats from synthetic inputs.
"""
from __future__ import annotations

import ctypes as C
import struct

import numpy as np

from . import lib


class rc_coarse_data(C.Structure):
    _fields_ = [("xdim", C.c_int), ("ydim", C.c_int), ("zdim", C.c_int), ("rmax", C.c_int),
                ("translation_num", C.c_int), ("orientation_num", C.c_int),
                ("image_size", C.c_int)] + [
        (n, C.POINTER(C.c_float)) for n in
        ("eulers", "trans_x", "trans_y", "real", "imag", "corr", "diff2s")]


def _bind():
    L = lib.load()
    L.rc_coarse_data_load.argtypes = [C.c_char_p, C.POINTER(rc_coarse_data)]
    L.rc_coarse_data_free.argtypes = [C.POINTER(rc_coarse_data)]
    L.rc_coarse_data_free.restype = None
    return L


def write_projector_dat(path, vol):
    """Projector dump: int32 xdim, ydim, zdim, inity, initz, maxr; float32
    padding; real plane; imaginary plane (acc_projector_impl.h:300-337)."""
    with open(path, "wb") as f:
        f.write(struct.pack("<6if", vol.mdlX, vol.mdlY, vol.mdlZ, vol.initY, vol.initZ,
                            vol.maxR, float(vol.pad)))
        f.write(np.ascontiguousarray(vol.real, "<f4").tobytes())
        f.write(np.ascontiguousarray(vol.imag, "<f4").tobytes())


def write_coarse_dat(path, xdim, ydim, zdim, rmax, eulers, trans_x, trans_y, real, imag, corr,
                     diff2s):
    """Coarse-particle dump in the layout main.cu:162-191 reads."""
    eulers = np.ascontiguousarray(eulers, "<f4").ravel()
    T, O, n = int(np.size(trans_x)), eulers.size // 9, int(np.size(real))
    with open(path, "wb") as f:
        f.write(struct.pack("<7i", xdim, ydim, zdim, rmax, T, O, n))
        for a in (eulers, trans_x, trans_y, real, imag, corr, diff2s):
            f.write(np.ascontiguousarray(a, "<f4").ravel().tobytes())


def load_coarse_dat(path):
    """Host arrays of a coarse-particle dump (through the C ABI loader)."""
    L = _bind()
    d = rc_coarse_data()
    lib.check(L.rc_coarse_data_load(path.encode(), C.byref(d)))
    try:
        O, T, n = d.orientation_num, d.translation_num, d.image_size
        arr = lambda p, k: np.ctypeslib.as_array(p, shape=(k,)).copy()
        return dict(xdim=d.xdim, ydim=d.ydim, zdim=d.zdim, rmax=d.rmax, translation_num=T,
                    orientation_num=O, image_size=n, eulers=arr(d.eulers, 9 * O),
                    trans_x=arr(d.trans_x, T), trans_y=arr(d.trans_y, T), real=arr(d.real, n),
                    imag=arr(d.imag, n), corr=arr(d.corr, n), diff2s=arr(d.diff2s, O * T))
    finally:
        L.rc_coarse_data_free(C.byref(d))


def replay(projector_path, data_path, alias_translations_to_eulers=False, do_CC=False):
    """Run one dump pair through rc_run_diff2_kernel_coarse; returns diff2s
    (initial values from the file plus the score, as the reference
    accumulates).  ``alias_translations_to_eulers`` reproduces the reference
    harness, which passes the Euler array as trans_x/trans_y
    (main.cu:234-235)."""
    import torch
    proj = lib.Projector(path=projector_path)
    d = load_coarse_dat(data_path)
    pk = proj.kernel(d["xdim"], d["ydim"], d["rmax"], imgZ=d["zdim"])
    t = lambda a: torch.as_tensor(a).cuda()
    eul = t(d["eulers"])
    T = d["translation_num"]
    if alias_translations_to_eulers:
        tx, ty = eul[:T], eul[T:2 * T]
    else:
        tx, ty = t(d["trans_x"]), t(d["trans_y"])
    out = t(d["diff2s"])
    lib.run_diff2_kernel_coarse(pk, tx, ty, t(d["corr"]), t(d["real"]), t(d["imag"]), eul, out,
                                d["orientation_num"], T, d["image_size"], do_CC=do_CC)
    torch.cuda.synchronize()
    res = out.cpu().numpy()
    proj.close()
    return res, d
