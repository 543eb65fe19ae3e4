"""Builds the CPU-oracle counterpart of a workload (test helper)."""
import pyoracle as O


def build_oracle(inp, nflux, ninterp, binned, theta0, exp_conf=0.0):
    bm = O.Beam(inp["beam"], inp["pixsize"], inp["minbeamval"],
                inp["nbeamhist"])
    L = O.Like(nflux, ninterp, inp["knots"])
    b = L.add_beamset(bm)
    if binned:
        bins, c0, dl = O.data_bin(inp["data"], inp["nbins"])
        L.add_field(b, bins=bins, bincent0=c0, bindelta=dl,
                    sigma_base=inp["sigma"], npix=inp["data"].size)
    else:
        L.add_field(b, pix=inp["data"], sigma_base=inp["sigma"])
    L.finalize(theta0, exp_conf=exp_conf)
    L._beam = bm
    return L
