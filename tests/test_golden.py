"""Golden fixtures (tests/golden/*.npz, written by make_golden.py with the
float64 oracle) stay reproducible: the generator regenerates the same
inputs and the oracle the same outputs (CPU only)."""
import os

import numpy as np

from oracle import oracle
from relion_amd import synth

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(HERE, name), allow_pickle=False)


def mini_problem_from_fixture(g):
    dims = g["vol_dims"]
    v = synth.ProjectorVolume(g["vol_real"], g["vol_imag"], int(dims[0]), int(dims[1]),
                              int(dims[2]), int(dims[3]), int(dims[4]), int(dims[5]),
                              float(g["pad"]))
    return v


def test_mini_fixture_inputs_regenerate():
    g = load("mini.npz")
    p = synth.make_problem("mini", seed=int(g["seed"]))
    np.testing.assert_array_equal(p.vols[0].real, g["vol_real"])
    np.testing.assert_array_equal(p.eulers, g["eulers"])
    np.testing.assert_array_equal(p.particles.real, g["real"])
    np.testing.assert_array_equal(p.particles.corr, g["corr"])


def test_mini_fixture_oracle_reproduces():
    g = load("mini.npz")
    v = mini_problem_from_fixture(g)
    imgY, maxR = (int(x) for x in g["img"])
    geom = synth.CoarseGeometry(imgY=imgY, maxR=maxR)
    off = g["trans_offset"]
    for i in range(g["real"].shape[0]):
        tx = g["trans_x"][off[i]:off[i + 1]]
        ty = g["trans_y"][off[i]:off[i + 1]]
        d = oracle.diff2_f64(v, geom.maxR2_padded, g["eulers"], tx, ty, g["real"][i],
                             g["imag"][i], g["corr"][i], g["coor_xy"][:, 0], g["coor_xy"][:, 1],
                             n_proj=int(g["npix"]))
        np.testing.assert_allclose(d, g["diff2"][i], rtol=1e-12)


def test_c1_digest():
    import hashlib
    g = load("c1_digest.npz")
    p = synth.make_problem("c1", seed=int(g["seed"]))
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(p.vols[0].real, np.float32).tobytes())
    h.update(np.ascontiguousarray(p.vols[0].imag, np.float32).tobytes())
    assert h.hexdigest() == str(g["volume_sha256"])
    d0 = oracle.problem_diff2(p, 0)
    np.testing.assert_allclose(d0, g["diff2_p0"], rtol=1e-12)
    for i in range(p.n_particles):
        d = oracle.problem_diff2(p, i)
        assert oracle.argmin_flat(d) == int(g["best_index"][i])
        assert abs(d.min() - g["best"][i]) <= 1e-12 * abs(g["best"][i])
