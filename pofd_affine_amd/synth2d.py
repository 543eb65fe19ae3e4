"""Synthetic two-band workloads for the 2D P(D) likelihood.

Deterministic from fixed seeds:

* counts: numberCountsDoubleLogNormal (include/numberCountsDoubleLogNormal.h)
  with band-1 knots at 3, 8, 20, 50, 120, 300 mJy taking the broken power
  law of the 1D workload, colour sigma knots (band-1 flux 5, 30, 200 mJy)
  with values 0.1, 0.2, 0.3 and colour offset knots at the same fluxes with
  values -0.2, -0.05, 0.1 (natural log of S2/S1);
* beams: circular Gaussians, FWHM 18.1" and 25.2", 6" pixels, both on the
  grid of the larger one cut at 9 FWHM, peak 1;
* map: sources Poisson-drawn over the band-1 knot range with the band-1
  counts, at uniformly random pixels, S2 = S1 exp(N(offset(S1), sigma(S1)))
  with the model's (piecewise-linear-between-knots here) colour law,
  periodically convolved with each beam, plus N(0, 2.5 / 2.7 mJy) noise,
  mean subtracted per band.
"""
import math

import numpy as np

from . import synth

KNOTS_MJY = (3.0, 8.0, 20.0, 50.0, 120.0, 300.0)
SIGMA_POS_MJY = (5.0, 30.0, 200.0)
SIGMA_VALS = (0.1, 0.2, 0.3)
OFFSET_POS_MJY = (5.0, 30.0, 200.0)
OFFSET_VALS = (-0.2, -0.05, 0.1)
FWHM1, FWHM2, PIX = 18.1, 25.2, 6.0
SIGMA1, SIGMA2 = 0.0025, 0.0027


def positions():
    """(knots, sigma knot positions, offset knot positions) in Jy."""
    return (np.array(KNOTS_MJY) * 1e-3, np.array(SIGMA_POS_MJY) * 1e-3,
            np.array(OFFSET_POS_MJY) * 1e-3)


def true_theta(sigma_mult=(1.0, 1.0), nbonus=4):
    """[log10 n1 at knots..., sigmas..., offsets..., sm1, sm2, bonus...]"""
    kv = np.log10(synth.dnds(np.array(KNOTS_MJY) * 1e-3))
    return np.concatenate([kv, SIGMA_VALS, OFFSET_VALS, sigma_mult,
                           np.zeros(nbonus)]).astype(np.float32)


def walker_thetas(nwalkers, seed=4321, scatter=0.05):
    """True values plus N(0, scatter) on the knot values and N(0, 0.01) on
    the colour parameters, one row per walker."""
    rng = np.random.default_rng(seed)
    base = true_theta().astype(np.float64)
    th = np.tile(base, (nwalkers, 1))
    nk, ns, no = len(KNOTS_MJY), len(SIGMA_VALS), len(OFFSET_VALS)
    th[:, :nk] += rng.normal(0.0, scatter, size=(nwalkers, nk))
    th[:, nk:nk + ns + no] += rng.normal(0.0, 0.01, size=(nwalkers, ns + no))
    return th.astype(np.float32)


def gaussian_beams(cut_fwhm=9.0):
    half = int(math.floor(0.5 * cut_fwhm * max(FWHM1, FWHM2) / PIX))
    y, x = np.mgrid[-half:half + 1, -half:half + 1].astype(np.float64)
    out = []
    for fw in (FWHM1, FWHM2):
        s = fw / (2.0 * math.sqrt(2.0 * math.log(2.0))) / PIX
        out.append(np.exp(-0.5 * (x * x + y * y) / (s * s)))
    return out[0], out[1], PIX


def _colour(s1, pos, vals):
    return np.interp(s1, np.array(pos) * 1e-3, vals)


def make_maps(npix_side=1000, seed=99, beams=None):
    """Synthetic pixel-pair map (two flattened float64 arrays), mean
    subtracted per band."""
    rng = np.random.default_rng(seed)
    if beams is None:
        b1, b2, pix = gaussian_beams()
    else:
        b1, b2, pix = beams
    smin, smax = KNOTS_MJY[0] * 1e-3, KNOTS_MJY[-1] * 1e-3
    area = (npix_side * pix / 3600.0) ** 2
    nsrc = rng.poisson(area * synth._n_between(smin, smax))
    # inverse-CDF draw restricted to [smin, smax]
    u = rng.random(nsrc)
    n_lo = synth._n_between(smin, synth.SBREAK)
    n_tot = synth._n_between(smin, smax)
    s1 = np.empty(nsrc)
    ulo = u * n_tot
    lo = ulo < n_lo
    g = synth.GAMMA1
    a = (smin / synth.S0) ** (1.0 - g)
    s1[lo] = synth.S0 * (a + ulo[lo] * (1.0 - g) / (synth.N0 * synth.S0)) ** (
        1.0 / (1.0 - g))
    g = synth.GAMMA2
    nb = synth.N0 * (synth.SBREAK / synth.S0) ** (-synth.GAMMA1)
    uh = ulo[~lo] - n_lo
    s1[~lo] = synth.SBREAK * (1.0 + uh * (1.0 - g) / (nb * synth.SBREAK)) ** (
        1.0 / (1.0 - g))
    mu = _colour(s1, OFFSET_POS_MJY, OFFSET_VALS)
    sg = _colour(s1, SIGMA_POS_MJY, SIGMA_VALS)
    s2 = s1 * np.exp(rng.normal(mu, sg))
    pos = rng.integers(0, npix_side * npix_side, size=nsrc)
    maps = []
    for flux, beam, sig in ((s1, b1, SIGMA1), (s2, b2, SIGMA2)):
        img = np.bincount(pos, weights=flux, minlength=npix_side ** 2).reshape(
            npix_side, npix_side)
        kern = np.zeros((npix_side, npix_side))
        h = beam.shape[0] // 2
        for dy in range(-h, h + 1):
            for dx in range(-h, h + 1):
                kern[dy % npix_side, dx % npix_side] = beam[dy + h, dx + h]
        conv = np.fft.irfft2(np.fft.rfft2(img) * np.fft.rfft2(kern),
                             s=img.shape)
        conv += rng.normal(0.0, sig, size=conv.shape)
        d = conv.ravel()
        maps.append(d - d.mean())
    return maps[0], maps[1], nsrc
