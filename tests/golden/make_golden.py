"""Writes the golden fixtures of this directory with the float64 oracle.

The reference repository holds no golden vectors for the coarse diff2 path
(its inputs are RELION dumps fetched from the network), so the fixtures are
produced by this build's own float64 oracle on seeded synthetic inputs and
committed, freezing the oracle's output so later changes to it, to the
generator or to the kernels are caught:

* mini.npz   -- a complete 16^3-map problem (volume, orientations,
                translations, 2 particles, weights) with its float64 diff2.
* c1_digest.npz -- the 64^3 / 32^2 / 576 x 21 x 8 workload: a digest of the
                volume, particle 0's full float64 diff2 and every particle's
                float64 min and argmin.

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "t2-relion-kernel_amd")]

from oracle import oracle  # noqa: E402
from relion_amd import synth  # noqa: E402

MINI_SEED = 31
C1_SEED = 11


def volume_digest(v):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(v.real, np.float32).tobytes())
    h.update(np.ascontiguousarray(v.imag, np.float32).tobytes())
    return h.hexdigest()


def main():
    p = synth.make_problem("mini", seed=MINI_SEED)
    v = p.vols[0]
    pb = p.particles
    diff2 = np.stack([oracle.problem_diff2(p, i) for i in range(p.n_particles)])
    np.savez_compressed(
        os.path.join(HERE, "mini.npz"), seed=MINI_SEED, vol_real=v.real, vol_imag=v.imag,
        vol_dims=np.array([v.mdlX, v.mdlY, v.mdlZ, v.initY, v.initZ, v.maxR], np.int32),
        pad=np.float32(v.pad), img=np.array([p.geom.imgY, p.geom.maxR], np.int32),
        eulers=p.eulers, trans_offset=p.trans_offset, trans_x=p.trans_x, trans_y=p.trans_y,
        real=pb.real, imag=pb.imag, corr=pb.corr, coor_xy=pb.coor_xy,
        npix=np.int32(pb.npix), diff2=diff2)
    c = synth.make_problem("c1", seed=C1_SEED)
    mins, args = [], []
    d0 = None
    for i in range(c.n_particles):
        d = oracle.problem_diff2(c, i)
        if i == 0:
            d0 = d
        mins.append(d.min())
        args.append(oracle.argmin_flat(d))
    np.savez_compressed(os.path.join(HERE, "c1_digest.npz"), seed=C1_SEED,
                        volume_sha256=volume_digest(c.vols[0]), diff2_p0=d0,
                        best=np.array(mins), best_index=np.array(args, np.int64))
    print("wrote mini.npz and c1_digest.npz")


if __name__ == "__main__":
    main()
