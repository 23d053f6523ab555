"""Seeded synthetic UPDC datasets (no WCS library needed).

Each exposure's detector-corner lattice (u, v) in {-0.5 .. iw-0.5} x
{-0.5 .. ih-0.5} is mapped straight to target pixel coordinates,

    T = s * R(theta) * [(u, v) + delta + D(u, v) - c] + c_T,

which is what the reference driver's vertex lattice holds after its WCS
round trip (updc/_engine.py:269-285).  Truth images, PSFs and noise follow the
workload descriptions the benchmark configurations name.  Observed data are
produced from a caller-supplied noise-free model (the GPU forward operator in
the product and bench, the oracle in CPU tests), so the generator itself does
no operator arithmetic.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

BASE_SEED = 20261203


def rng_for(config: int, stream: str) -> np.random.Generator:
    sub = {"truth": 1, "geometry": 2, "psf": 3, "noise": 4}[stream]
    return np.random.Generator(np.random.PCG64([BASE_SEED + config, sub]))


@dataclass
class Geometry:
    iw: int
    ih: int
    scale: float = 2.0
    theta_deg: float = 0.0
    dither: tuple[float, float] = (0.0, 0.0)
    cubic: np.ndarray | None = None      # (2, 10) coefficients of the cubic distortion, detector px
    center_target: tuple[float, float] = (0.0, 0.0)

    def vertices(self) -> np.ndarray:
        u = np.arange(self.iw + 1, dtype=np.float64) - 0.5
        v = np.arange(self.ih + 1, dtype=np.float64) - 0.5
        uu, vv = np.meshgrid(u, v)
        cu, cv = (self.iw - 1) / 2.0, (self.ih - 1) / 2.0
        du = uu + self.dither[0]
        dv = vv + self.dither[1]
        if self.cubic is not None:
            a = uu / (self.iw / 2.0) - 1.0
            b = vv / (self.ih / 2.0) - 1.0
            terms = [np.ones_like(a), a, b, a * a, a * b, b * b, a ** 3, a * a * b, a * b * b, b ** 3]
            du = du + sum(c * t for c, t in zip(self.cubic[0], terms))
            dv = dv + sum(c * t for c, t in zip(self.cubic[1], terms))
        th = np.deg2rad(self.theta_deg)
        ct, st = np.cos(th), np.sin(th)
        x = du - cu
        y = dv - cv
        tx = self.scale * (ct * x - st * y) + self.center_target[0]
        ty = self.scale * (st * x + ct * y) + self.center_target[1]
        return np.ascontiguousarray(np.stack([tx, ty], axis=-1))


@dataclass
class Dataset:
    W: int
    H: int
    geoms: list[Geometry]
    psfs: list[np.ndarray]
    truth: np.ndarray
    background: list[float]
    read_noise: float
    mask_fraction: float = 0.0
    pedestal: list[float] | None = None          # None = automatic rule
    iterations: int = 20
    name: str = ""
    sci: list[np.ndarray] = field(default_factory=list)
    wht: list[np.ndarray] = field(default_factory=list)

    def vertices(self, k: int) -> np.ndarray:
        return self.geoms[k].vertices()


# ---------------------------------------------------------------- sources
def gaussian_psf(size: int, fwhm: float, dx: float = 0.0, dy: float = 0.0) -> np.ndarray:
    s = fwhm / (2.0 * np.sqrt(2.0 * np.log(2.0)))
    c = (size - 1) / 2.0
    y, x = np.mgrid[0:size, 0:size]
    p = np.exp(-((x - c - dx) ** 2 + (y - c - dy) ** 2) / (2 * s * s))
    return (p / p.sum()).astype(np.float32)


def airy_wing_psf(size: int, fwhm: float, wing: float = 3e-3, rng: np.random.Generator | None = None) -> np.ndarray:
    """Airy core (2 J1(x)/x)^2 with the given FWHM plus a 1/r^3 halo."""
    from scipy.special import j1
    c = (size - 1) / 2.0
    y, x = np.mgrid[0:size, 0:size].astype(np.float64)
    r = np.hypot(x - c, y - c)
    lam_d = fwhm / 1.028993969962188  # FWHM of the Airy pattern = 1.029 lambda/D
    z = np.pi * r / lam_d
    with np.errstate(invalid="ignore", divide="ignore"):
        core = np.where(z > 1e-8, (2.0 * j1(z) / np.where(z > 1e-8, z, 1.0)) ** 2, 1.0)
    halo = wing / (1.0 + (r / fwhm) ** 2) ** 1.5
    p = core + halo
    return (p / p.sum()).astype(np.float32)


def _render(img: np.ndarray, x0: float, y0: float, flux: float, prof, radius: int) -> None:
    H, W = img.shape
    ix0, iy0 = int(np.floor(x0)) - radius, int(np.floor(y0)) - radius
    xs0, ys0 = max(ix0, 0), max(iy0, 0)
    xs1, ys1 = min(ix0 + 2 * radius + 2, W), min(iy0 + 2 * radius + 2, H)
    if xs0 >= xs1 or ys0 >= ys1:
        return
    y, x = np.mgrid[ys0:ys1, xs0:xs1].astype(np.float64)
    stamp = prof(x - x0, y - y0)
    tot = stamp.sum()
    if tot > 0:
        img[ys0:ys1, xs0:xs1] += flux * stamp / tot


def truth_image(W: int, H: int, nsrc: int, sky: float, rng: np.random.Generator, sersic_fraction: float = 0.0,
                sigma_range=(0.5, 3.0), flux_mu=6.0, flux_sigma=1.5) -> np.ndarray:
    img = np.full((H, W), sky, np.float64)
    xs = rng.uniform(8, W - 8, nsrc)
    ys = rng.uniform(8, H - 8, nsrc)
    fl = rng.lognormal(flux_mu, flux_sigma, nsrc)
    kind = rng.uniform(0, 1, nsrc) < sersic_fraction
    sig = rng.uniform(*sigma_range, nsrc)
    nser = rng.uniform(0.5, 4.0, nsrc)
    for i in range(nsrc):
        if kind[i]:
            re, n = sig[i] * 1.5, nser[i]
            bn = 2.0 * n - 1.0 / 3.0 + 0.009876 / n
            prof = (lambda dx, dy, re=re, n=n, bn=bn:
                    np.exp(-bn * ((np.hypot(dx, dy) / re) ** (1.0 / n) - 1.0)))
            radius = int(min(64, max(6, 8 * re)))
        else:
            s = sig[i]
            prof = (lambda dx, dy, s=s: np.exp(-(dx * dx + dy * dy) / (2 * s * s)))
            radius = int(max(4, 5 * s))
        _render(img, xs[i], ys[i], fl[i], prof, radius)
    return img


# ---------------------------------------------------------------- configurations
def config1() -> Dataset:
    """512^2 target, 200 Gaussians (sigma U[0.5,3], lognormal flux 6/1.5),
    sky 0.1; N=4 detectors of 256^2 at 2x scale, dithers U[0,1), rotation
    U[-0.5,0.5] deg; Gaussian PSF FWHM 2 in 63^2; read noise 5; pedestal 50;
    20 iterations."""
    g = rng_for(1, "geometry")
    geoms = [Geometry(256, 256, 2.0, float(g.uniform(-0.5, 0.5)), (float(g.uniform(0, 1)), float(g.uniform(0, 1))),
                      None, (255.5, 255.5)) for _ in range(4)]
    truth = truth_image(512, 512, 200, 0.1, rng_for(1, "truth"))
    psfs = [gaussian_psf(63, 2.0) for _ in range(4)]
    return Dataset(512, 512, geoms, psfs, truth, [0.0] * 4, 5.0, 0.0, [50.0] * 4, 20, "config1")


def config2(n_exp: int = 8, det: int = 2048, target: int = 4096, psf_size: int = 1431, nsrc: int = 5000) -> Dataset:
    """N=8 detectors of 2048^2, affine + cubic distortion (coeffs N(0,1e-3)),
    dithers U[0,10]; 4096^2 target; Airy+1/r^3 PSF (FWHM 2.2) of 1431^2 per
    exposure; 5000 Gaussian/Sersic sources + sky 0.3; read noise 10; ~1% of
    pixels masked; 40 iterations."""
    g = rng_for(2, "geometry")
    geoms = []
    c = (target - 1) / 2.0
    for _ in range(n_exp):
        th = float(g.uniform(-0.5, 0.5))
        d = (float(g.uniform(0, 10)), float(g.uniform(0, 10)))
        cub = g.normal(0.0, 1e-3, (2, 10))
        geoms.append(Geometry(det, det, target / det, th, d, cub, (c, c)))
    pr = rng_for(2, "psf")
    psfs = [airy_wing_psf(psf_size, 2.2 * float(pr.uniform(0.98, 1.02))) for _ in range(n_exp)]
    truth = truth_image(target, target, nsrc, 0.3, rng_for(2, "truth"), sersic_fraction=0.5)
    return Dataset(target, target, geoms, psfs, truth, [0.0] * n_exp, 10.0, 0.01, None, 40, "config2")


def config3() -> Dataset:
    """Low-sky case of the config-2 workload: the same geometry and PSFs,
    sky 0.01.  Background offsets cannot start pixels at raw_mu <= 0: the
    initial estimate subtracts the weighted mean background, so raw_mu(x0) =
    H_k x0 + ped_k + (b_k - mean b), and offsets large enough to matter make
    the signed multiplicative update diverge.  FreezeMu is forced instead by
    the reference's own recipe (every background -100)."""
    ds = config2()
    ds.truth = truth_image(ds.W, ds.H, 5000, 0.01, rng_for(3, "truth"), sersic_fraction=0.5)
    ds.name = "config3"
    return ds


def config3_stress() -> Dataset:
    """FreezeMu stress on the config-3 workload: sky 0.01 and no read noise.
    A detector pixel then sees 0.04 sky counts, so about 96% of sky pixels
    record 0, the automatic pedestal is 1e-12 (_engine.py:368-372 with sigma
    = 0), and the drizzled estimate is 0 wherever every overlapping pixel
    recorded 0: there raw_mu = H_k x + b_k + ped_k starts at the pedestal,
    and whether it stays above 0 is decided by transform rounding."""
    ds = config3()
    ds.read_noise = 0.0
    ds.name = "config3_stress"
    return ds


def gaussian_mixture_psf(size: int, rng: np.random.Generator, ncomp: int = 3,
                         fwhm_range=(1.8, 3.0)) -> np.ndarray:
    """Normalised sum of `ncomp` Gaussians, FWHM U[fwhm_range], weights
    U[0.2, 1], centroid offsets N(0, 0.3) px."""
    c = (size - 1) / 2.0
    y, x = np.mgrid[0:size, 0:size].astype(np.float64)
    p = np.zeros((size, size))
    for _ in range(ncomp):
        s = float(rng.uniform(*fwhm_range)) / (2.0 * np.sqrt(2.0 * np.log(2.0)))
        dx, dy = rng.normal(0.0, 0.3, 2)
        p += float(rng.uniform(0.2, 1.0)) * np.exp(-((x - c - dx) ** 2 + (y - c - dy) ** 2) / (2 * s * s))
    return (p / p.sum()).astype(np.float32)


def config4(target: int = 6144, nsrc: int = 11250) -> Dataset:
    """Uneven batching: N=24 exposures in one launch -- 12 of 2048^2, 8
    subarrays of 1024 (wide) x 2048 (high) and 4 of 640^2 -- at 2x scale on a
    6144^2 target, dithers U[0,40] px, rotations U[-2,2] deg; each with its
    own 3-component Gaussian-mixture PSF (FWHM U[1.8,3] px) of 1431^2 (even
    exposures) or 511^2 (odd).  Sources scale with the area of config 2."""
    g = rng_for(4, "geometry")
    sizes = [(2048, 2048)] * 12 + [(1024, 2048)] * 8 + [(640, 640)] * 4
    geoms = []
    for i, (iw, ih) in enumerate(sizes):
        fw, fh = 2 * iw, 2 * ih  # footprint in target pixels
        # full detectors tile the target in a 2x2 pattern; subarrays land anywhere inside it
        if i < 12:
            cx = fw / 2 + (target - fw) * (i % 2)
            cy = fh / 2 + (target - fh) * ((i // 2) % 2)
        else:
            cx = float(g.uniform(fw / 2, target - fw / 2))
            cy = float(g.uniform(fh / 2, target - fh / 2))
        th = float(g.uniform(-2.0, 2.0))
        d = (float(g.uniform(0, 40)), float(g.uniform(0, 40)))
        geoms.append(Geometry(iw, ih, 2.0, th, d, None, (cx - 0.5 - 40.0, cy - 0.5 - 40.0)))
    pr = rng_for(4, "psf")
    psfs = [gaussian_mixture_psf(1431 if i % 2 == 0 else 511, pr) for i in range(len(sizes))]
    truth = truth_image(target, target, nsrc, 0.3, rng_for(4, "truth"), sersic_fraction=0.5)
    return Dataset(target, target, geoms, psfs, truth, [0.0] * len(sizes), 10.0, 0.01, None, 40, "config4")


def config5_psf_params(n: int = 5) -> np.ndarray:
    """(n, 2, 4) true PSF mixture parameters (a, s, dx, dy) of config 5:
    a core of sigma ~0.9 and a wider component of sigma ~2, amplitude 0.2."""
    g = rng_for(5, "psf")
    p = np.zeros((n, 2, 4))
    for k in range(n):
        p[k, 0] = [1.0, float(g.uniform(0.8, 1.0)), float(g.normal(0, 0.1)), float(g.normal(0, 0.1))]
        p[k, 1] = [0.2, float(g.uniform(1.8, 2.4)), float(g.normal(0, 0.2)), float(g.normal(0, 0.2))]
    return p


def mixture_image(params: np.ndarray, size: int) -> np.ndarray:
    """Normalised float32 image of a (ncomp, 4) mixture, the same formula as
    updc_amd.autograd.gaussian_mixture."""
    c = (size - 1) / 2.0
    ax = np.arange(size, dtype=np.float64) - c
    y, x = ax[:, None], ax[None, :]
    img = sum(a * np.exp(-((x - dx) ** 2 + (y - dy) ** 2) / (2.0 * s * s)) for a, s, dx, dy in params)
    return (img / img.sum()).astype(np.float32)


def config5(target: int = 2048, det: int = 1024, psf_size: int = 255, nsrc: int = 1250) -> Dataset:
    """Differentiation workload: N=4 exposures of 1024^2 plus a held-out fifth
    at 2x scale on a 2048^2 target, 255^2 PSFs that are 2-component Gaussian
    mixtures (config5_psf_params), 10 unrolled iterations."""
    g = rng_for(5, "geometry")
    c = (target - 1) / 2.0
    geoms = [Geometry(det, det, target / det, float(g.uniform(-0.5, 0.5)),
                      (float(g.uniform(0, 10)), float(g.uniform(0, 10))), g.normal(0.0, 1e-3, (2, 10)), (c, c))
             for _ in range(5)]
    psfs = [mixture_image(p, psf_size) for p in config5_psf_params(5)]
    truth = truth_image(target, target, nsrc, 0.3, rng_for(5, "truth"), sersic_fraction=0.5)
    return Dataset(target, target, geoms, psfs, truth, [0.0] * 5, 10.0, 0.01, None, 10, "config5")


def small_config(n_exp=2, det=48, target=96, psf=9, seed=7, rotation=3.0, dither=3.0, distort=0.0,
                 nsrc=20, sky=0.5, mask_fraction=0.02, read_noise=2.0) -> Dataset:
    """Tiny dataset for oracle-sized parity tests."""
    g = np.random.Generator(np.random.PCG64([BASE_SEED, seed]))
    c = (target - 1) / 2.0
    geoms = []
    for _ in range(n_exp):
        cub = g.normal(0, distort, (2, 10)) if distort > 0 else None
        geoms.append(Geometry(det, det, target / det, float(g.uniform(-rotation, rotation)),
                              (float(g.uniform(0, dither)), float(g.uniform(0, dither))), cub, (c, c)))
    truth = truth_image(target, target, nsrc, sky, g, sersic_fraction=0.3, sigma_range=(0.6, 2.0))
    psfs = [gaussian_psf(psf, 2.0 + 0.2 * k, 0.1 * k, -0.05 * k) for k in range(n_exp)]
    return Dataset(target, target, geoms, psfs, truth, [0.0] * n_exp, read_noise, mask_fraction, None, 5,
                   f"small{seed}")


# ---------------------------------------------------------------- observation
def observe(ds: Dataset, model: list[np.ndarray], config_seed: int = 0) -> Dataset:
    """Poisson(model) + N(0, rn^2) data, W = 1/variance, a masked fraction
    with W = 0, and the pedestal (given, or the reference's automatic rule
    max(0, max(1e-12, 1e-6 sigma) - min SCI), updc/_engine.py:368-372)."""
    rng = rng_for(config_seed, "noise") if config_seed else np.random.Generator(np.random.PCG64([BASE_SEED, 99]))
    ds.sci, ds.wht = [], []
    peds = []
    for k, m in enumerate(model):
        lam = np.clip(m, 0, None)
        y = rng.poisson(lam).astype(np.float64) + rng.normal(0.0, ds.read_noise, m.shape)
        var = lam + ds.read_noise ** 2
        with np.errstate(divide="ignore"):
            w = (1.0 / var).astype(np.float32)  # inf where var = 0: masked by the weight rule
        if ds.mask_fraction > 0:
            w[rng.uniform(0, 1, m.shape) < ds.mask_fraction] = 0.0
        sci = y.astype(np.float32)
        ds.sci.append(sci)
        ds.wht.append(w)
        if ds.pedestal is None:
            sigma = ds.read_noise
            floor = max(1e-12, 1e-6 * sigma)
            valid = w > 0
            peds.append(float(max(0.0, floor - float(np.min(sci[valid].astype(np.float64))))))
    if ds.pedestal is None:
        ds.pedestal = peds
    return ds
