"""GPU parity of the prep kernel's layouts: the knot spline at knot counts
other than the usual eight, and beams past the LDS stage.

The prep kernel fits the knot spline on one wavefront, a lane per knot
(numberCountsKnotsSpline::setParams and GSL's natural cubic spline,
src/numberCountsKnotsSpline.cc:265-274), so the counts at the edges of that
layout -- two knots (no system to solve), three (a one-equation system) and
64 (every lane) -- and two in between are checked against the CPU oracle:
R at the interpolation nodes within 1e-12 (log2 R) and the log likelihood
within 1e-6 of the exact-math value (the oracle with extended-precision
transforms: with two knots the model fits the map badly, the data sit in
the P(D)'s far tail, and a plain double transform is off by ~1e-6 there),
on the small 200x200 map at nflux 2^15.
"""
import numpy as np
import pytest

from pofd_affine_amd import synth, workload
from oracle_side import build_oracle, check_exact_parity

pytestmark = pytest.mark.gpu

NF, NI = 32768, 512
NWALK = 4


@pytest.mark.parametrize("nk", [2, 3, 5, 17, 64])
def test_knot_counts(nk):
    print("[gpu-case] knot count %d" % nk, flush=True)
    inp = dict(workload.make_inputs(npix_side=200, seed=7))
    knots = np.geomspace(synth.knot_positions()[0], synth.knot_positions()[-1], nk)
    inp["knots"] = knots
    rng = np.random.default_rng(100 + nk)
    base = np.concatenate([np.log10(synth.dnds(knots)), [1.0, 0.0, 0.0]])
    th = np.tile(base, (NWALK, 1))
    th[:, :nk] += rng.normal(0.0, 0.03, size=(NWALK, nk))
    th = th.astype(np.float32)
    ctx = workload.build_context(inp, NF, NI, binned=True)
    L = build_oracle(inp, NF, NI, True, th[0])
    ll, st = ctx.loglike_batch(th)
    assert np.all(st == 0), st
    check_exact_parity(ctx, L, th, ll, tol=1e-6, nthreads=NWALK)
    L.set_precise(0)
    for w in range(NWALK):
        ip = L.initpd(0, th[w])
        x, y = ctx.get_rinterp(w)
        np.testing.assert_array_equal(x, ip["rinterp_flux"])
        np.testing.assert_allclose(y, ip["rinterp_vals"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("nhist", [0, 120])
def test_raw_beam_past_the_lds_stage(nhist):
    """A raw (not histogrammed) beam of more entries than the prep kernel
    stages in LDS (1,024; a Gaussian beam at 2" pixels has ~1,300 above
    1e-6), so R at the nodes reads the beam from global memory; and the same
    beam histogrammed, read from LDS.  Node values and logL against the
    oracle as above."""
    import pyoracle as O
    print("[gpu-case] raw beam, nhist %d" % nhist, flush=True)
    inp = dict(workload.make_inputs(npix_side=200, seed=7))
    beam2d, pix = synth.gaussian_beam(pix_arcsec=2.0)
    inp.update(beam=beam2d.ravel(), pixsize=pix, nbeamhist=nhist)
    ob = O.Beam(inp["beam"], pix, inp["minbeamval"], nhist)
    npos = int(ob.info()["npos"])
    if nhist == 0:
        assert npos > 1024, npos
    th = synth.walker_thetas(NWALK, seed=3)
    ctx = workload.build_context(inp, NF, NI, binned=True)
    L = build_oracle(inp, NF, NI, True, th[0])
    ll, st = ctx.loglike_batch(th)
    assert np.all(st == 0), st
    check_exact_parity(ctx, L, th, ll, tol=1e-6, nthreads=NWALK)
    L.set_precise(0)
    for w in range(NWALK):
        ip = L.initpd(0, th[w])
        x, y = ctx.get_rinterp(w)
        np.testing.assert_array_equal(x, ip["rinterp_flux"])
        np.testing.assert_allclose(y, ip["rinterp_vals"], rtol=0, atol=1e-12)
