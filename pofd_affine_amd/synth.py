"""Synthetic single-band workloads for the 1D P(D) likelihood.

The reference ships no benchmark inputs, so the measured configurations are
generated here, deterministically from fixed seeds:

* counts: a broken power law, dN/dS = 5e6 (S/10 mJy)^-2.2 Jy^-1 deg^-2 below
  20 mJy and continuing with slope -3.5 above, sampled at knots
  2, 5, 10, 20, 45, 100, 200, 450 mJy (log10 values are the "true" theta of
  a numberCountsKnotsSpline model, include/numberCountsKnotsSpline.h);
* beam: circular Gaussian, FWHM 18.1", 6" pixels, cut at 9 FWHM
  (2*floor(4.5*FWHM/pix)+1 pixels on a side), peak 1;
* map: Poisson-drawn sources over 1-600 mJy at uniformly random pixels,
  periodically convolved with the beam, plus N(0, sigma) noise, mean
  subtracted (what fitsData::meanSubtract does, src/fitsData.cc:393-401).
"""
import math

import numpy as np

KNOTS_MJY = (2.0, 5.0, 10.0, 20.0, 45.0, 100.0, 200.0, 450.0)
N0 = 5e6          # dN/dS at S0, Jy^-1 deg^-2
S0 = 0.010        # Jy
SBREAK = 0.020    # Jy
GAMMA1 = 2.2
GAMMA2 = 3.5
SMIN_SRC = 0.001  # Jy
SMAX_SRC = 0.600  # Jy


def dnds(s):
    """Broken power-law differential counts (Jy^-1 deg^-2)."""
    s = np.asarray(s, dtype=np.float64)
    lo = N0 * (s / S0) ** (-GAMMA1)
    nb = N0 * (SBREAK / S0) ** (-GAMMA1)
    hi = nb * (s / SBREAK) ** (-GAMMA2)
    return np.where(s <= SBREAK, lo, hi)


def knot_positions():
    return np.array(KNOTS_MJY, dtype=np.float64) * 1e-3


def true_theta(sigma_mult=1.0, nbonus=2):
    """[log10 dN/dS at the knots..., sigma_mult, bonus slots...] (float32)."""
    vals = np.log10(dnds(knot_positions()))
    return np.concatenate([vals, [sigma_mult], np.zeros(nbonus)]).astype(
        np.float32)


def walker_thetas(nwalkers, seed=1234, scatter=0.05, sigma_mult=1.0,
                  nbonus=2):
    """True knots plus N(0, scatter) in log10, one row per walker."""
    rng = np.random.default_rng(seed)
    base = true_theta(sigma_mult, nbonus).astype(np.float64)
    th = np.tile(base, (nwalkers, 1))
    nk = len(KNOTS_MJY)
    th[:, :nk] += rng.normal(0.0, scatter, size=(nwalkers, nk))
    return th.astype(np.float32)


def gaussian_beam(fwhm_arcsec=18.1, pix_arcsec=6.0, cut_fwhm=9.0):
    """Peak-normalised pixelised Gaussian beam (2-D array) and pixel size."""
    half = int(math.floor(0.5 * cut_fwhm * fwhm_arcsec / pix_arcsec))
    side = 2 * half + 1
    sig = fwhm_arcsec / (2.0 * math.sqrt(2.0 * math.log(2.0))) / pix_arcsec
    y, x = np.mgrid[-half:half + 1, -half:half + 1].astype(np.float64)
    b = np.exp(-0.5 * (x * x + y * y) / (sig * sig))
    assert b.shape == (side, side)
    return b, pix_arcsec


def _n_between(a, b):
    """Integral of dnds over [a, b] (closed form per segment)."""
    def seg(lo, hi, n_at, s_at, g):
        if hi <= lo:
            return 0.0
        return n_at * s_at / (1.0 - g) * ((hi / s_at) ** (1.0 - g)
                                          - (lo / s_at) ** (1.0 - g))
    nb = N0 * (SBREAK / S0) ** (-GAMMA1)
    return (seg(a, min(b, SBREAK), N0, S0, GAMMA1)
            + seg(max(a, SBREAK), b, nb, SBREAK, GAMMA2))


def _draw_fluxes(rng, n, smax=SMAX_SRC):
    """Inverse-CDF draw of n fluxes in [SMIN_SRC, smax]."""
    n_lo = _n_between(SMIN_SRC, SBREAK)
    n_tot = _n_between(SMIN_SRC, smax)
    u = rng.random(n) * n_tot
    out = np.empty(n)
    lo = u < n_lo
    # segment 1: N(S) = N0 S0/(1-g) [(S/S0)^(1-g) - (Smin/S0)^(1-g)]
    g = GAMMA1
    a = (SMIN_SRC / S0) ** (1.0 - g)
    out[lo] = S0 * (a + u[lo] * (1.0 - g) / (N0 * S0)) ** (1.0 / (1.0 - g))
    g = GAMMA2
    nb = N0 * (SBREAK / S0) ** (-GAMMA1)
    a = 1.0
    uh = u[~lo] - n_lo
    out[~lo] = SBREAK * (a + uh * (1.0 - g) / (nb * SBREAK)) ** (1.0 / (1.0 - g))
    return out


def make_map(npix_side=1000, sigma_noise=0.0025, seed=42, beam=None,
             pix_arcsec=6.0, smax=SMAX_SRC):
    """Synthetic confused map (flattened float64), mean subtracted.

    smax: brightest source drawn (Jy); the measured configurations use
    600 mJy, above the last knot (450 mJy), which puts the brightest pixels
    in the far tail of the model P(D)."""
    rng = np.random.default_rng(seed)
    if beam is None:
        beam, pix_arcsec = gaussian_beam(pix_arcsec=pix_arcsec)
    area_deg2 = (npix_side * pix_arcsec / 3600.0) ** 2
    nsrc = rng.poisson(area_deg2 * _n_between(SMIN_SRC, smax))
    flux = _draw_fluxes(rng, nsrc, smax)
    pos = rng.integers(0, npix_side * npix_side, size=nsrc)
    img = np.bincount(pos, weights=flux,
                      minlength=npix_side * npix_side).reshape(
                          npix_side, npix_side)
    # periodic convolution with the beam centred at (0, 0)
    kern = np.zeros((npix_side, npix_side))
    h = beam.shape[0] // 2
    for dy in range(-h, h + 1):
        for dx in range(-h, h + 1):
            kern[dy % npix_side, dx % npix_side] = beam[dy + h, dx + h]
    conv = np.fft.irfft2(np.fft.rfft2(img) * np.fft.rfft2(kern),
                         s=img.shape)
    conv += rng.normal(0.0, sigma_noise, size=conv.shape)
    d = conv.ravel()
    d = d - d.mean()
    return d, nsrc
