"""In-radius partition of coarse images (FilterCoarseParamImage,
main.cu:88-151) through the host entry point rc_filter_coarse_order (CPU
only: it reads the view's geometry and touches no device)."""
import numpy as np
import pytest

from relion_amd import lib, synth


def _coords(imgX, imgY, maxR, n):
    i = np.arange(n)
    x, y = i % imgX, i // imgX
    y = np.where(y > maxR, y - imgY, y)
    return x, y


@pytest.mark.parametrize("imgY,count", [(32, 414), (64, 1636), (96, 3654)])
def test_inside_count_is_the_half_disc_lattice_count(imgY, count):
    """pad 2 and maxR = imgY/2: maxR2_padded = 4 maxR^2 exactly, so the
    inside pixels are the lattice points x >= 0 of the disc x^2 + y^2 <=
    maxR^2 in the stored rows (the 414 / 1636 / 3654 of the workloads)."""
    imgX, maxR = imgY // 2 + 1, imgY // 2
    n = imgX * imgY
    order, n_in = lib.filter_order(lib.geometry_view(imgX, imgY, maxR, 2.0), n)
    assert n_in == count
    x, y = _coords(imgX, imgY, maxR, n)
    disc = x * x + y * y <= maxR * maxR
    assert n_in == int(disc.sum())
    # head: the disc's pixels in index order; tail: the others, last first
    assert np.array_equal(order[:n_in], np.flatnonzero(disc))
    assert np.array_equal(order[n_in:], np.flatnonzero(~disc)[::-1])
    assert np.array_equal(np.sort(order), np.arange(n))


def test_matches_the_test_data_generator_order():
    for imgY in (16, 32, 64):
        g = synth.CoarseGeometry(imgY=imgY, maxR=imgY // 2)
        ref, n_ref = synth.filter_order(g)
        order, n_in = lib.filter_order(lib.geometry_view(g.imgX, g.imgY, g.maxR, g.pad),
                                       g.image_size)
        assert n_in == n_ref and np.array_equal(order, ref)


def test_float_rule_with_a_non_integer_padding_factor():
    """pad 1.7: r2 = float(float(x^2 + y^2) * pad) * pad against the float of
    maxR2_padded.  The integer rule x^2 + y^2 <= maxR^2 differs from it on
    the boundary pixels (float rounding), and the entry point must follow
    the float rule, pixel by pixel."""
    imgY, maxR, pad = 40, 17, np.float32(1.7)
    imgX = imgY // 2 + 1
    n = imgX * imgY
    m2p = int(maxR * maxR * pad * pad)             # makeKernel's truncation
    order, n_in = lib.filter_order(lib.geometry_view(imgX, imgY, maxR, float(pad), m2p), n)
    x, y = _coords(imgX, imgY, maxR, n)
    r2 = (x * x + y * y).astype(np.float32) * pad * pad      # float32 at each step
    inside = r2 <= np.float32(m2p)
    assert n_in == int(inside.sum())
    assert np.array_equal(order[:n_in], np.flatnonzero(inside))
    assert np.array_equal(order[n_in:], np.flatnonzero(~inside)[::-1])
    # the float rule is not the integer one here: some boundary pixel differs
    assert np.any(inside != (x * x + y * y <= maxR * maxR))


def test_partial_image_and_rejections():
    """image_size below imgX*imgY covers the leading pixels only; a size
    beyond the image is an error."""
    imgY, maxR = 32, 16
    imgX = 17
    order, n_in = lib.filter_order(lib.geometry_view(imgX, imgY, maxR, 2.0), 100)
    assert np.array_equal(np.sort(order), np.arange(100))
    x, y = _coords(imgX, imgY, maxR, 100)
    assert n_in == int((x * x + y * y <= maxR * maxR).sum())
    with pytest.raises(lib.RcError):
        lib.filter_order(lib.geometry_view(imgX, imgY, maxR, 2.0), imgX * imgY + 1)
