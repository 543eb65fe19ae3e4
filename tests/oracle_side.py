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


def check_exact_parity(ctx, L, th, ll, tol=1e-6, nthreads=16, max_off=8):
    """Device log likelihoods `ll` of a one-beam-set, one-field context
    against the extended-precision oracle L (the exact-math value of the
    reference's algorithm), within tol per walker.

    The split search's argmin (src/PDFactory.cc:783-795) is ill-conditioned
    where two neighbouring P(D) values differ by less than a double
    transform's rounding there; one index moves logL by ~3e-6 at 10^6
    pixels.  A walker beyond tol must be such a case: its split one index
    from the exact one, the two candidates' exact values closer than the
    double-precision oracle's own error there, and the exact value at the
    device's split within tol.  Returns the walkers so handled."""
    import numpy as np
    L.set_precise(1)
    ext, st_e = L.loglike_batch(th, nthreads)
    assert np.all(st_e == 0), st_e
    off = np.nonzero(np.abs(ll - ext) > tol)[0]
    assert off.size <= max_off, (off, ll[off] - ext[off])
    sk = L.nknots
    for w in off:
        w = int(w)
        gsplit = int(ctx.get_debug(w)["split"])
        L.set_precise(1)
        L.initpd(0, th[w])
        r_ext = L.getpd(0, 0, float(th[w][sk]), setlog=False)
        L.set_precise(0)
        L.initpd(0, th[w])
        r_dbl = L.getpd(0, 0, float(th[w][sk]), setlog=False)
        esplit = r_ext["split"]
        assert abs(gsplit - esplit) == 1, (w, gsplit, esplit)
        pe, pdb = r_ext["pofd_raw"], r_dbl["pofd_raw"]
        lo, hi = max(0, esplit - 64), esplit + 64
        noise = np.max(np.abs(pdb[lo:hi] - pe[lo:hi]))
        assert abs(pe[gsplit] - pe[esplit]) < noise, (w, pe[gsplit],
                                                      pe[esplit], noise)
        L.set_precise(1)
        L.set_split_shift(gsplit - esplit)
        at_dev, st_a = L.loglike_batch(th[w:w + 1], 1)
        L.set_split_shift(0)
        assert st_a[0] == 0
        assert abs(ll[w] - at_dev[0]) <= tol, (w, ll[w] - at_dev[0])
    L.set_precise(0)
    return off
