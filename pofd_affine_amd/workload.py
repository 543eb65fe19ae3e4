"""Builds the measured 1D workloads on a device context.

Configuration names follow the measured setups: ``config1d`` is the batched
1D fit of a 1000x1000-pixel synthetic map at nflux = 2^17 with a
120-bin histogrammed Gaussian beam (FWHM 18.1", 6" pixels, values < 1e-6
dropped) and sigma_inst = 2.5 mJy; smaller sides/nflux give the quick
parity cases.
"""
import numpy as np

from . import synth
from ._lib import Beam, Context, data_bin


def make_inputs(npix_side=1000, sigma=0.0025, seed=42, minbeamval=1e-6,
                nbeamhist=120, nbins=1000, smax=synth.SMAX_SRC):
    beam2d, pixsize = synth.gaussian_beam()
    data, nsrc = synth.make_map(npix_side, sigma, seed=seed, smax=smax)
    return dict(beam=beam2d.ravel(), pixsize=pixsize, data=data, nsrc=nsrc,
                sigma=sigma, minbeamval=minbeamval, nbeamhist=nbeamhist,
                nbins=nbins, knots=synth.knot_positions())


def build_context(inp, nflux=131072, ninterp=2048, binned=True, device=0,
                  exp_conf=0.0, spectral_cut=None):
    """A prepared Context with one beam set and one field."""
    ctx = Context(device)
    if spectral_cut is not None:
        ctx.set_spectral_cut(spectral_cut)
    ctx.configure(nflux, ninterp, inp["knots"], exp_conf)
    beam = Beam(inp["beam"], inp["pixsize"], inp["minbeamval"],
                inp["nbeamhist"])
    b = ctx.add_beamset(beam)
    if binned:
        bins, c0, dl = data_bin(inp["data"], inp["nbins"])
        ctx.add_field(b, bins=bins, bincent0=c0, bindelta=dl,
                      sigma_base=inp["sigma"], npix=inp["data"].size)
    else:
        ctx.add_field(b, pix=inp["data"], sigma_base=inp["sigma"])
    ctx.prepare()
    ctx._beam = beam
    return ctx


def uneven_fields(seed=2024, nfield=8):
    """The multi-field workload: nfield maps of 2e5-4e6 pixels (log-uniform
    sizes), noise 1.5-4 mJy and nflux 2^15-2^18, as (side, sigma, nflux,
    map seed) tuples."""
    rng = np.random.default_rng(seed)
    npix = np.exp(rng.uniform(np.log(2e5), np.log(4e6), nfield))
    sides = np.maximum(64, np.round(np.sqrt(npix)).astype(int))
    sigmas = rng.uniform(0.0015, 0.004, nfield)
    nfluxes = (2 ** rng.integers(15, 19, nfield)).astype(int)
    return [(int(sides[i]), float(sigmas[i]), int(nfluxes[i]), 300 + i)
            for i in range(nfield)]
