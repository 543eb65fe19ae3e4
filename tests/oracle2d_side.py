"""Builds the 2D CPU oracle (test infrastructure) for a synthetic two-band
workload, the same inputs the device path gets."""
import numpy as np

import pyoracle2d as O2
from pofd_affine_amd import synth2d


def make_inputs2d(npix_side=150, seed=99, nbeamhist=120, minbeamval=1e-6,
                  nbins=None):
    b1, b2, pix = synth2d.gaussian_beams()
    d1, d2, nsrc = synth2d.make_maps(npix_side, seed, (b1, b2, pix))
    out = dict(beam1=b1, beam2=b2, pixsize=pix, d1=d1, d2=d2, nsrc=nsrc,
               nbeamhist=nbeamhist, minbeamval=minbeamval, nbins=nbins)
    if nbins:
        bins, info = O2.data2_bin(d1, d2, nbins, nbins)
        out["bins"], out["bininfo"] = bins, info
    return out


def build_oracle2d(inp, n, theta, nedge=256, binned=False, edge_integ=True,
                   sigma=(synth2d.SIGMA1, synth2d.SIGMA2)):
    L = O2.Like2(n, nedge, edge_integ)
    kp, sp, op = synth2d.positions()
    L.set_positions(kp, sp, op)
    bm = O2.DoubleBeam(inp["beam1"], inp["beam2"], inp["pixsize"],
                       inp["minbeamval"], inp["nbeamhist"])
    b = L.add_beamset(bm)
    if binned:
        L.add_field(b, inp["d1"], inp["d2"], sigma[0], sigma[1],
                    bins=inp["bins"], bininfo=inp["bininfo"])
    else:
        L.add_field(b, inp["d1"], inp["d2"], sigma[0], sigma[1])
    L.finalize(theta)
    return L
