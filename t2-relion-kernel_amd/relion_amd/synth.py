"""Seeded synthetic inputs for the coarse diff2 search.

Everything a caller of the coarse-search entry points hands over is built here
from a seed: a Gaussian-blob density map, its zero-padded half-complex Fourier
volume in RELION's projector layout, a HEALPix x psi orientation set, an
integer Gauss-circle translation set and CTF-modulated noisy particle images
with per-pixel weights.

Layout conventions (all follow the reference's data contract):

* Fourier volume: planar float32 ``real``/``imag`` arrays of ``mdlX*mdlY*mdlZ``
  with x fastest, then y, then z (acc_projector_impl.h:186-196); x holds only
  the non-negative half (kx = 0..mdlX-1), y and z are centred with
  ``initY = initZ = -(pad_size // 2)`` (the padded-size rule RELION uses when
  it builds ``AccProjector`` objects).
* Orientation ``o`` is 9 float32 entries, a row-major 3x3 matrix ``M``; only
  ``e0,e1,e3,e4,e6,e7`` are read, ``(xp, yp, zp) = pad * M @ (x, y, 0)``
  (acc_projectorkernel_impl.h:178-192).
* Translations are SoA ``trans_x[T]``, ``trans_y[T]`` already scaled so the
  phase is ``x*tx + y*ty`` radians (cuda_device_utils.cuh:110-129).
* A coarse image of ``imgY`` rows has ``imgX = imgY//2 + 1`` columns; pixel
  ``i`` sits at ``x = i % imgX, y = i // imgX`` with ``y -= imgY`` when
  ``y > maxR`` (diff2.cuh:182-188).

The data generator is not the checker: particles are made with its own
float64 trilinear forward model, independent of the test oracle.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "ProjectorVolume", "CoarseGeometry", "ParticleBatch", "CoarseProblem",
    "blob_map", "blob_params", "fourier_volume", "fourier_volume_analytic",
    "healpix_centres", "euler_zyz",
    "orientation_set", "translation_set", "pixel_coords", "filter_order",
    "project_np", "make_particles", "make_problem", "CONFIGS",
]


# --------------------------------------------------------------------------
# density map and Fourier volume
# --------------------------------------------------------------------------
def blob_map(n: int, n_blobs: int = 20, seed: int = 0,
             sigma_range=(2.0, 6.0), centre_radius: float | None = None) -> np.ndarray:
    """Sum of ``n_blobs`` isotropic Gaussians on an ``n**3`` grid (float64).

    Centres are uniform in a ball of radius ``centre_radius`` voxels (default
    ``20 * n / 64``) around the box centre ``n/2``; sigmas ``U(sigma_range)``
    voxels; amplitudes ``U(0.5, 1)``.
    """
    grid = np.arange(n, dtype=np.float64)
    out = np.zeros((n, n, n), dtype=np.float64)
    for c, s, a in blob_params(n, n_blobs, seed, sigma_range, centre_radius):
        gx = np.exp(-(grid - c[0]) ** 2 / (2 * s * s))
        gy = np.exp(-(grid - c[1]) ** 2 / (2 * s * s))
        gz = np.exp(-(grid - c[2]) ** 2 / (2 * s * s))
        out += a * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    return out


@dataclass
class ProjectorVolume:
    """Padded half-complex Fourier volume in the projector's memory layout."""
    real: np.ndarray          # float32 [mdlZ*mdlY*mdlX]
    imag: np.ndarray          # float32 [mdlZ*mdlY*mdlX]
    mdlX: int
    mdlY: int
    mdlZ: int
    initY: int
    initZ: int
    maxR: int                 # model radius in unpadded pixels
    pad: float

    @property
    def shape(self):
        return (self.mdlZ, self.mdlY, self.mdlX)


def fourier_volume(density: np.ndarray, pad: int = 2, scale: float | None = None) -> ProjectorVolume:
    """Zero-pad ``density`` ``pad``-fold, rFFT it and lay the half volume out
    the way ``AccProjector`` stores it.

    The map centre is moved to the origin before the transform so Fourier
    coefficients vary smoothly (what trilinear interpolation assumes).  The
    volume keeps ``|k| <= pad * r_max`` with ``r_max = n/2`` and zeroes the
    rest, ``pad_size = 2*(pad*r_max + 1) + 1``.
    """
    n = density.shape[0]
    p = pad * n
    big = np.zeros((p, p, p), dtype=np.float64)
    o = (p - n) // 2
    big[o:o + n, o:o + n, o:o + n] = density
    f = np.fft.rfftn(np.fft.ifftshift(big))            # [kz, ky, kx], kx in [0, p/2]
    r_max = n // 2
    pad_size = 2 * (pad * r_max + 1) + 1
    mdlX = pad_size // 2 + 1
    init = -(pad_size // 2)
    kyz = np.arange(pad_size) + init
    kx = np.arange(mdlX)
    kx_ok = kx <= p // 2
    vol = np.zeros((pad_size, pad_size, mdlX), dtype=np.complex128)
    vol[:, :, kx_ok] = f[np.ix_(kyz % p, kyz % p, kx[kx_ok])]
    kz3 = kyz[:, None, None]
    ky3 = kyz[None, :, None]
    kx3 = kx[None, None, :]
    vol[(kx3 ** 2 + ky3 ** 2 + kz3 ** 2) > (pad * r_max) ** 2] = 0
    if scale is None:
        scale = 1.0 / float(n) ** 1.5
    vol *= scale
    return ProjectorVolume(
        real=np.ascontiguousarray(vol.real.astype(np.float32).ravel()),
        imag=np.ascontiguousarray(vol.imag.astype(np.float32).ravel()),
        mdlX=mdlX, mdlY=pad_size, mdlZ=pad_size, initY=init, initZ=init,
        maxR=r_max, pad=float(pad))


def blob_params(n: int, n_blobs: int = 20, seed: int = 0, sigma_range=(2.0, 6.0),
                centre_radius: float | None = None):
    """The (centre, sigma, amplitude) triples ``blob_map`` draws for a seed."""
    rng = np.random.default_rng(seed)
    if centre_radius is None:
        centre_radius = 20.0 * n / 64.0
    out = []
    for _ in range(n_blobs):
        while True:
            c = rng.uniform(-1.0, 1.0, 3)
            if np.dot(c, c) <= 1.0:
                break
        c = n / 2.0 + centre_radius * c
        s = rng.uniform(*sigma_range)
        a = rng.uniform(0.5, 1.0)
        out.append((c, s, a))
    return out


def fourier_volume_analytic(n: int, r_max: int, pad: int = 2, seed: int = 0,
                            n_blobs: int = 20) -> ProjectorVolume:
    """Projector volume of the ``blob_map(n, seed=seed)`` map computed from
    the closed-form Fourier transform of its Gaussians, only out to radius
    ``r_max`` (so a coarse-pass volume of a large box costs no padded FFT):
    F(k) = sum_b a_b (2 pi s_b^2)^{3/2} exp(-2 pi^2 s_b^2 |k|^2 / P^2)
           exp(-2 pi i k . (c_b - n/2) / P),  P = pad * n,
    scaled like ``fourier_volume``."""
    p = pad * n
    pad_size = 2 * (pad * r_max + 1) + 1
    mdlX = pad_size // 2 + 1
    init = -(pad_size // 2)
    kyz = np.arange(pad_size, dtype=np.float64) + init
    kx = np.arange(mdlX, dtype=np.float64)
    kz3, ky3, kx3 = kyz[:, None, None], kyz[None, :, None], kx[None, None, :]
    k2 = kx3 ** 2 + ky3 ** 2 + kz3 ** 2
    vol = np.zeros((pad_size, pad_size, mdlX), dtype=np.complex128)
    for c, s, a in blob_params(n, n_blobs, seed):
        d = c - n / 2.0
        amp = a * (2 * math.pi * s * s) ** 1.5 * np.exp(-2 * math.pi ** 2 * s * s * k2 / (p * p))
        ph = np.exp(-2j * math.pi * d[0] * kx / p)[None, None, :] * \
            np.exp(-2j * math.pi * d[1] * kyz / p)[None, :, None] * \
            np.exp(-2j * math.pi * d[2] * kyz / p)[:, None, None]
        vol += amp * ph
    vol[k2 > (pad * r_max) ** 2] = 0
    vol *= 1.0 / float(n) ** 1.5
    return ProjectorVolume(
        real=np.ascontiguousarray(vol.real.astype(np.float32).ravel()),
        imag=np.ascontiguousarray(vol.imag.astype(np.float32).ravel()),
        mdlX=mdlX, mdlY=pad_size, mdlZ=pad_size, initY=init, initZ=init,
        maxR=r_max, pad=float(pad))


# --------------------------------------------------------------------------
# orientations and translations
# --------------------------------------------------------------------------
def healpix_centres(order: int):
    """Pixel centres (theta, phi) of the HEALPix RING scheme, Nside = 2**order.

    Closed form of Gorski et al. (2005), ApJ 622:759, section 4 (polar caps
    and equatorial belt).
    """
    ns = 1 << order
    npix = 12 * ns * ns
    ncap = 2 * ns * (ns - 1)
    z = np.empty(npix)
    phi = np.empty(npix)
    for p in range(npix):
        if p < ncap:                                  # north polar cap
            ph = (p + 1) / 2.0
            i = int(math.floor(math.sqrt(ph - math.sqrt(math.floor(ph))))) + 1
            j = p + 1 - 2 * i * (i - 1)
            z[p] = 1.0 - i * i / (3.0 * ns * ns)
            phi[p] = math.pi / (2.0 * i) * (j - 0.5)
        elif p < npix - ncap:                         # equatorial belt
            pp = p - ncap
            i = pp // (4 * ns) + ns
            j = pp % (4 * ns) + 1
            s = (i - ns + 1) % 2
            z[p] = 4.0 / 3.0 - 2.0 * i / (3.0 * ns)
            phi[p] = math.pi / (2.0 * ns) * (j - s / 2.0)
        else:                                         # south polar cap
            pp = npix - p
            ph = pp / 2.0
            i = int(math.floor(math.sqrt(ph - math.sqrt(math.floor(ph))))) + 1
            j = 4 * i + 1 - (pp - 2 * i * (i - 1))
            z[p] = -1.0 + i * i / (3.0 * ns * ns)
            phi[p] = math.pi / (2.0 * i) * (j - 0.5)
    return np.arccos(np.clip(z, -1.0, 1.0)), phi


def euler_zyz(rot: float, tilt: float, psi: float) -> np.ndarray:
    """Rotation matrix of ZYZ Euler angles (radians), the convention RELION
    and Xmipp use for (rot, tilt, psi): ``A = Rz(psi) Ry(tilt) Rz(rot)``."""
    ca, sa = math.cos(rot), math.sin(rot)
    cb, sb = math.cos(tilt), math.sin(tilt)
    cg, sg = math.cos(psi), math.sin(psi)
    cc, cs, sc, ss = cb * ca, cb * sa, sb * ca, sb * sa
    return np.array([
        [cg * cc - sg * sa, cg * cs + sg * ca, -cg * sb],
        [-sg * cc - cg * sa, -sg * cs + cg * ca, sg * sb],
        [sc, ss, cb]])


def orientation_set(order: int, n_psi: int) -> np.ndarray:
    """``12*4**order * n_psi`` orientations as float32 ``[O, 9]``.

    Each row is the inverse (transpose) of the ZYZ matrix of
    (rot=phi, tilt=theta, psi=k*360/n_psi): it maps the image plane into the
    volume, which is what the projector's ``e0..e8`` hold.
    """
    theta, phi = healpix_centres(order)
    rows = []
    for t, f in zip(theta, phi):
        for k in range(n_psi):
            rows.append(euler_zyz(f, t, 2.0 * math.pi * k / n_psi).T.ravel())
    return np.asarray(rows, dtype=np.float32)


def translation_set(r2_max: int, box: int, centre=(0.0, 0.0)):
    """Integer shifts ``(dx, dy)`` with ``dx**2 + dy**2 <= r2_max`` around
    ``centre`` (pixels of a ``box``-pixel particle), as phase slopes
    ``tx = 2*pi*(cx+dx)/box`` (float32 SoA)."""
    r = int(math.isqrt(r2_max))
    d = [(dx, dy) for dy in range(-r, r + 1) for dx in range(-r, r + 1)
         if dx * dx + dy * dy <= r2_max]
    d = np.asarray(d, dtype=np.float64)
    sx = d[:, 0] + centre[0]
    sy = d[:, 1] + centre[1]
    tx = (2.0 * math.pi * sx / box).astype(np.float32)
    ty = (2.0 * math.pi * sy / box).astype(np.float32)
    return tx, ty, np.stack([sx, sy], 1)


# --------------------------------------------------------------------------
# image geometry
# --------------------------------------------------------------------------
@dataclass
class CoarseGeometry:
    imgY: int                  # coarse image rows
    maxR: int                  # image radius (pixels)
    pad: float = 2.0

    @property
    def imgX(self) -> int:
        return self.imgY // 2 + 1

    @property
    def image_size(self) -> int:
        return self.imgX * self.imgY

    @property
    def maxR2_padded(self) -> int:
        return int(self.maxR * self.maxR * self.pad * self.pad)


def pixel_coords(g: CoarseGeometry):
    """Half-plane pixel index -> integer (x, y) (diff2.cuh:182-188)."""
    i = np.arange(g.image_size)
    x = i % g.imgX
    y = i // g.imgX
    y = np.where(y > g.maxR, y - g.imgY, y)
    return x.astype(np.int64), y.astype(np.int64)


def filter_order(g: CoarseGeometry):
    """Permutation putting in-radius pixels first (stable) and out-of-radius
    pixels reversed at the tail, plus the in-radius count (main.cu:88-151)."""
    x, y = pixel_coords(g)
    inside = (x * x + y * y) * g.pad * g.pad <= g.maxR2_padded
    head = np.nonzero(inside)[0]
    tail = np.nonzero(~inside)[0][::-1]
    return np.concatenate([head, tail]), int(head.size)


# --------------------------------------------------------------------------
# forward model for particle generation (float64)
# --------------------------------------------------------------------------
def _trilinear_clamped(vol3: np.ndarray, xp, yp, zp):
    nz, ny, nx = vol3.shape
    x0 = np.floor(xp).astype(np.int64)
    y0 = np.floor(yp).astype(np.int64)
    z0 = np.floor(zp).astype(np.int64)
    fx, fy, fz = xp - x0, yp - y0, zp - z0
    xa, xb = np.clip(x0, 0, nx - 1), np.clip(x0 + 1, 0, nx - 1)
    ya, yb = np.clip(y0, 0, ny - 1), np.clip(y0 + 1, 0, ny - 1)
    za, zb = np.clip(z0, 0, nz - 1), np.clip(z0 + 1, 0, nz - 1)
    c = lambda zz, yy, xx: vol3[zz, yy, xx]
    c00 = c(za, ya, xa) * (1 - fx) + c(za, ya, xb) * fx
    c01 = c(za, yb, xa) * (1 - fx) + c(za, yb, xb) * fx
    c10 = c(zb, ya, xa) * (1 - fx) + c(zb, ya, xb) * fx
    c11 = c(zb, yb, xa) * (1 - fx) + c(zb, yb, xb) * fx
    c0 = c00 * (1 - fy) + c01 * fy
    c1 = c10 * (1 - fy) + c11 * fy
    return c0 * (1 - fz) + c1 * fz


def _complex_volume(v: ProjectorVolume) -> np.ndarray:
    c = getattr(v, "_c128", None)
    if c is None:
        c = (v.real.astype(np.float64) + 1j * v.imag.astype(np.float64)).reshape(v.shape)
        object.__setattr__(v, "_c128", c)
    return c


def project_np(v: ProjectorVolume, m9: np.ndarray, x: np.ndarray, y: np.ndarray,
               maxR2_padded: int) -> np.ndarray:
    """Central-slice sample of the half volume at image pixels (x, y) for the
    3x3 matrix ``m9`` (float64 throughout, exact trilinear, clamp-to-edge,
    Hermitian pairing for negative rotated x)."""
    vol3 = _complex_volume(v)
    m = np.asarray(m9, dtype=np.float64).reshape(3, 3)
    xp = (m[0, 0] * x + m[0, 1] * y) * v.pad
    yp = (m[1, 0] * x + m[1, 1] * y) * v.pad
    zp = (m[2, 0] * x + m[2, 1] * y) * v.pad
    r2 = xp * xp + yp * yp + zp * zp
    neg = xp < 0
    sgn = np.where(neg, -1.0, 1.0)
    val = _trilinear_clamped(vol3, xp * sgn, yp * sgn - v.initY, zp * sgn - v.initZ)
    val = np.where(neg, np.conj(val), val)
    return np.where(np.floor(r2) <= maxR2_padded, val, 0.0)


def _random_rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def _ctf(s2: np.ndarray, defocus: float, wavelength: float = 0.0197, q0: float = 0.1):
    chi = math.pi * wavelength * defocus * s2
    return -(math.sqrt(1 - q0 * q0) * np.sin(chi) + q0 * np.cos(chi))


@dataclass
class ParticleBatch:
    """Particle images in the filtered pixel order (in-radius first)."""
    real: np.ndarray          # float32 [P, full_npix]
    imag: np.ndarray          # float32 [P, full_npix]
    corr: np.ndarray          # float32 [P, full_npix]  per-pixel weight w
    coor_xy: np.ndarray       # float32 [full_npix, 2] integer-valued (x, y)
    npix: int                 # in-radius pixel count (the first npix entries)
    full_npix: int
    true_rot: np.ndarray      # float64 [P, 3, 3] generating orientation
    true_shift: np.ndarray    # float64 [P, 2] generating shift (pixels)
    classes: np.ndarray       # int32 [P] generating class
    snr: float

    @property
    def n(self) -> int:
        return self.real.shape[0]


def make_particles(vols, g: CoarseGeometry, n: int, box: int, snr: float,
                   seed: int, shift_range: float = 2.0, ctf_thresh: float = 0.05,
                   apix: float = 1.5, defocus_range=(5000.0, 30000.0),
                   centres=None, noise=True) -> ParticleBatch:
    """Particles = CTF x (projection at a random continuous orientation,
    shifted by a random sub-pixel offset) + complex Gaussian noise at ``snr``
    (signal power over noise power on in-radius pixels).

    The weights are ``w = CTF^2 / sigma^2`` and the image handed over is
    ``X / CTF`` (zero where ``|CTF| < ctf_thresh``), so
    ``sum w/2 |P - X/CTF|^2 = sum |CTF P - X|^2 / (2 sigma^2)``.
    ``centres`` (optional ``[n, 2]``) offsets the shift draw per particle.
    """
    if isinstance(vols, ProjectorVolume):
        vols = [vols]
    rng = np.random.default_rng(seed)
    order, npix = filter_order(g)
    x, y = pixel_coords(g)
    x, y = x[order].astype(np.float64), y[order].astype(np.float64)
    full = g.image_size
    s2 = (x * x + y * y) / float(box * apix) ** 2
    real = np.zeros((n, full), np.float32)
    imag = np.zeros((n, full), np.float32)
    corr = np.zeros((n, full), np.float32)
    rots = np.zeros((n, 3, 3))
    shifts = np.zeros((n, 2))
    cls = np.zeros(n, np.int32)
    for i in range(n):
        c = int(rng.integers(len(vols)))
        rot = _random_rotation(rng)
        sh = rng.uniform(-shift_range, shift_range, 2)
        if centres is not None:
            sh = sh + np.asarray(centres[i], dtype=np.float64)
        defocus = rng.uniform(*defocus_range)
        p = project_np(vols[c], rot.ravel(), x, y, g.maxR2_padded)
        ctf = _ctf(s2, defocus)
        sig = ctf * p * np.exp(-2j * math.pi * (x * sh[0] + y * sh[1]) / box)
        pw = float(np.mean(np.abs(sig[:npix]) ** 2))
        sigma2 = pw / snr if pw > 0 else 1.0
        obs = sig
        if noise:
            obs = sig + rng.normal(scale=math.sqrt(sigma2 / 2), size=full) \
                + 1j * rng.normal(scale=math.sqrt(sigma2 / 2), size=full)
        ok = np.abs(ctf) >= ctf_thresh
        xt = np.where(ok, obs / np.where(ok, ctf, 1.0), 0.0)
        w = np.where(ok, ctf * ctf / sigma2, 0.0)
        real[i] = xt.real
        imag[i] = xt.imag
        corr[i] = w
        rots[i] = rot
        shifts[i] = sh
        cls[i] = c
    coor = np.stack([x, y], 1).astype(np.float32)
    return ParticleBatch(real, imag, corr, coor, npix, full, rots, shifts, cls, snr)


# --------------------------------------------------------------------------
# named workloads
# --------------------------------------------------------------------------
@dataclass
class CoarseProblem:
    name: str
    vols: list
    geom: CoarseGeometry
    eulers: np.ndarray            # float32 [O, 9]
    trans_offset: np.ndarray      # int32 [P+1] CSR into trans_x/trans_y
    trans_x: np.ndarray           # float32 [sum T]
    trans_y: np.ndarray
    particles: ParticleBatch
    box: int
    meta: dict = field(default_factory=dict)

    @property
    def n_particles(self) -> int:
        return self.particles.n

    @property
    def n_classes(self) -> int:
        return len(self.vols)

    @property
    def orientation_num(self) -> int:
        return self.eulers.shape[0]

    def trans_of(self, i: int):
        a, b = int(self.trans_offset[i]), int(self.trans_offset[i + 1])
        return self.trans_x[a:b], self.trans_y[a:b]

    def evals(self) -> int:
        t = np.diff(self.trans_offset).astype(np.int64)
        return int(t.sum()) * self.orientation_num * self.n_classes


# name -> (map n, coarse imgY, healpix order, n_psi, r2_max, particles, snr, shift range)
CONFIGS = {
    "mini": dict(n=16, imgY=16, order=0, n_psi=4, r2=2, particles=2, snr=0.5, shift=1.0,
                 classes=1),
    "c1": dict(n=64, imgY=32, order=1, n_psi=12, r2=5, particles=8, snr=0.1, shift=2.0,
               classes=1),
    "c2": dict(n=128, imgY=64, order=2, n_psi=24, r2=36, particles=512, snr=0.05, shift=2.0,
               classes=1),
    # batched 3D classification: per-particle translation radius 4..8 px
    # around a random prior centre (49..197 shifts each)
    "c4": dict(n=128, imgY=64, order=2, n_psi=24, r2=None, radius=(4, 8), particles=2048,
               snr=0.05, shift=2.0, classes=8),
}


def make_problem(name: str = "c1", seed: int = 1234, particles: int | None = None,
                 classes: int | None = None, noise: bool = True,
                 snr: float | None = None) -> CoarseProblem:
    """Build one of the named workloads (``mini``, ``c1``, ``c2``, ``c4``).

    ``c1``: 64^3 map, 32^2 coarse image (414 in-radius pixels), 576
    orientations (HEALPix order 1 x 12 psi), 21 shifts (r^2 <= 5), 8 particles
    at SNR 0.1.  ``c2``: 128^3 map, 64^2 coarse image (1636 in-radius
    pixels), 4608 orientations (order 2 x 24 psi), 113 shifts (r^2 <= 36),
    512 particles at SNR 0.05.  ``c4``: 8 independent 128^3 classes, 2048
    particles, each with its own translation set of radius 4-8 px around a
    random prior centre (49-197 shifts).  ``mini``: a 16^3 map for fixtures.
    """
    c = CONFIGS[name]
    n_cls = c["classes"] if classes is None else classes
    vols = [fourier_volume(blob_map(c["n"], seed=seed + 7919 * k)) for k in range(n_cls)]
    g = CoarseGeometry(imgY=c["imgY"], maxR=c["imgY"] // 2, pad=2.0)
    eul = orientation_set(c["order"], c["n_psi"])
    n = c["particles"] if particles is None else particles
    rng = np.random.default_rng(seed + 2)
    if c["r2"] is not None:
        tx, ty, _ = translation_set(c["r2"], c["n"])
        off = (np.arange(n + 1) * tx.size).astype(np.int32)
        txs, tys = np.tile(tx, n), np.tile(ty, n)
        centres = None
    else:
        centres = rng.uniform(-2.0, 2.0, (n, 2))
        txl, tyl, counts = [], [], []
        for i in range(n):
            r = int(rng.integers(c["radius"][0], c["radius"][1] + 1))
            tx, ty, _ = translation_set(r * r, c["n"], centre=tuple(centres[i]))
            txl.append(tx)
            tyl.append(ty)
            counts.append(tx.size)
        off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        txs, tys = np.concatenate(txl), np.concatenate(tyl)
    parts = make_particles(vols, g, n, c["n"], c["snr"] if snr is None else snr, seed + 1,
                           shift_range=c["shift"], noise=noise, centres=centres)
    return CoarseProblem(name, vols, g, eul, off, txs, tys, parts, c["n"],
                         meta=dict(seed=seed, **c))
