"""The C ABI library loads, exports every symbol the header declares, and
rejects bad calls with status codes before touching a device (CPU only)."""
import ctypes as C
import os
import re
import struct

import pytest

from relion_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "relion_coarse.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(rc_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    L = lib.load()
    names = header_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n
    assert sorted(lib.EXPORTS) == names


def test_version_and_variants():
    L = lib.load()
    assert b"gfx950" in L.rc_version()
    for form in ("fp32", "bf16x3", "fp16"):
        v = lib.variants(form)
        assert len(v) >= 3
        for _, tm, tn in v:
            assert tm % 16 == 0 and tn % 16 == 0
    with pytest.raises(lib.RcError):
        lib.check(L.rc_variant_info(0, 99, None, None, None))


def test_default_variant_indices_name_their_kernels():
    """The host picks defaults by table index: 14 (wide 256 x 256,
    staggered), 21 (wide 128 x 512, fenced), 23 (wide 256 x 256, operand
    precomputed), 24 (direct 128 x 512) and 25 (direct 128 x 512, operand
    precomputed) must be those kernels."""
    for form in ("fp32", "bf16x3", "fp16"):
        assert "256x256" in lib.variant_name(form, 14) and "staggered" in lib.variant_name(form, 14)
        assert "128x512" in lib.variant_name(form, 21) and "fenced" in lib.variant_name(form, 21)
        assert "256x256" in lib.variant_name(form, 23)
        assert "precomputed" in lib.variant_name(form, 23)
        assert lib.variant_name(form, 24).startswith("cross_kernel_direct")
        assert "128x512" in lib.variant_name(form, 24)
        assert "precomputed" not in lib.variant_name(form, 24)
        assert lib.variant_name(form, 25).startswith("cross_kernel_direct")
        assert "precomputed" in lib.variant_name(form, 25)


def test_workspace_query():
    L = lib.load()
    a = L.rc_coarse_workspace_bytes(113, 4608, 1636, 2112, 0)
    b = L.rc_coarse_workspace_bytes(113, 4608, 3654, 4704, 0)
    assert 0 < a < b
    assert a % 256 == 0
    assert L.rc_coarse_workspace_bytes(1, 1, 0, 0, 0) == 0


def test_make_kernel_follows_reference_rule():
    """maxR = min(model maxr, image maxR), maxR2_padded = int(maxR^2 pad^2)
    (acc_projectorkernel_impl.h:58, 410-433) -- exercised on a host-only
    struct through the error path and through the public rule."""
    L = lib.load()
    k = lib.rc_projector_kernel()
    assert L.rc_projector_make_kernel(None, 17, 32, 1, 16, C.byref(k)) == 1
    assert b"NULL" in L.rc_last_error()


def test_error_paths_without_device(tmp_path):
    L = lib.load()
    h = C.c_void_p()
    st = L.rc_projector_create_from_file(str(tmp_path / "missing.dat").encode(), C.byref(h))
    assert st == 6 and not h.value
    trunc = tmp_path / "trunc.dat"
    trunc.write_bytes(struct.pack("<6if", 66, 131, 131, -65, -65, 32, 2.0) + b"\0" * 64)
    assert L.rc_projector_create_from_file(str(trunc).encode(), C.byref(h)) == 6
    assert b"truncated" in L.rc_last_error()
    two_d = tmp_path / "twod.dat"
    two_d.write_bytes(struct.pack("<6if", 4, 4, 1, -2, 0, 2, 2.0) + b"\0" * 128)
    # zdim == 1 is a 2D reference: not on this path
    buf = (C.c_float * 16)()
    assert L.rc_projector_create(4, 4, 1, -2, 0, 2, 2.0, buf, buf, C.byref(h)) == 2
    # 3D data is rejected before anything runs
    k = lib.rc_projector_kernel()
    st = L.rc_run_diff2_kernel_coarse(C.byref(k), None, None, None, None, None, None, None,
                                      None, 1.0, 1, 1, 1, None, 0, 1)
    assert st == 2 and b"data_is_3D" in L.rc_last_error()
    # empty plan description
    d = lib.rc_batch_desc()
    p = C.c_void_p()
    assert L.rc_plan_create(C.byref(d), C.byref(p)) == 1
    assert L.rc_plan_create(None, C.byref(p)) == 1


def test_single_run_validates_before_launch():
    L = lib.load()
    k = lib.rc_projector_kernel()
    # NULL projector view
    st = L.rc_coarse_run(21, 576, 414, 544, 0, None, None, None, None, None, C.byref(k), None,
                         None, None, None, None, None, 0)
    assert st == 1


def test_new_entry_points_validate_before_launch():
    """The split query, the partition calls and the allocation-free drop-in
    reject bad arguments with status codes, touching no device."""
    L = lib.load()
    ks, kc = C.c_int(), C.c_int()
    assert L.rc_coarse_split(0, 4608, 1636, 0, C.byref(ks), C.byref(kc)) == 1
    assert L.rc_coarse_split(113, 4608, 1636, 9, C.byref(ks), C.byref(kc)) == 1
    assert L.rc_coarse_split(113, 4608, 1636, 1, C.byref(ks), C.byref(kc)) == 0
    assert ks.value > 1 and (ks.value - 1) * kc.value < 52 <= ks.value * kc.value
    # one slice when the image alone fills the device or has one chunk
    assert L.rc_coarse_split(113, 4608 * 64, 1636, 1, C.byref(ks), C.byref(kc)) == 0
    assert ks.value == 1
    assert L.rc_coarse_split(21, 576, 20, 1, C.byref(ks), C.byref(kc)) == 0
    assert ks.value == 1 and kc.value == 1
    view = lib.geometry_view(17, 32, 16, 2.0)
    n = C.c_int()
    assert L.rc_filter_coarse_order(None, 544, None, C.byref(n)) == 1
    assert L.rc_filter_coarse_order(C.byref(view), 0, None, C.byref(n)) == 1
    assert L.rc_filter_coarse_images(C.byref(view), 544, 1, None, None, None, None, None, None,
                                     None, None) == 1
    assert L.rc_filter_coarse_images(C.byref(view), 545 + 17 * 32, 1, None, None, None, None,
                                     None, None, None, None) == 1
    k = lib.rc_projector_kernel()
    st = L.rc_run_diff2_kernel_coarse_noalloc(C.byref(k), None, None, None, None, None, None,
                                              None, None, 1.0, 1, 1, 1, None, 0, 1)
    assert st == 2 and b"data_is_3D" in L.rc_last_error()
    st = L.rc_run_diff2_kernel_coarse_noalloc(C.byref(k), None, None, None, None, None, None,
                                              None, None, 1.0, 1, 1, 1, None, 0, 0)
    assert st == 1


def test_c_caller_example_is_built():
    """The C-ABI-only example caller is built with the library and finds it
    next to itself (rpath $ORIGIN/..)."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "t2-relion-kernel_amd", "examples", "coarse_search")
    assert os.access(exe, os.X_OK), exe
    out = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    line = [ln for ln in out.splitlines() if "librelion_coarse_amd" in ln]
    assert line and "not found" not in line[0], out

