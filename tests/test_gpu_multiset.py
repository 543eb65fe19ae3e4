"""GPU parity with several beam sets and several datasets in one context.

calcLike groups its datasets by beam (src/calcLike.cc:688-758) and sums
(logL - like_offset) / like_norm over every dataset of every beam set
(calcLikeSingle::getLogLike :389-443, calcLike::getLogLike :912-916).  On
the device that is the per-dataset noise rows of the forward kernel, the
inverse kernel's dataset index, the dataset sum of the combine kernel and
the accumulation across beam-set launches; these cases compare all of them
with the oracle at the 1e-6 bar: two beam sets (a positive Gaussian beam
and the reference's two-signed test beam) with two datasets each, of
different noise, one binned and one unbinned, one with like_norm != 1.
The two-band case does the same with two beam sets of two datasets.
"""
import os

import numpy as np
import pytest

import pyoracle as O
import pyoracle2d as O2
from conftest import REFDATA
from oracle2d_side import make_inputs2d
from pofd_affine_amd import _lib, fitsio, synth, synth2d

pytestmark = pytest.mark.gpu

NF, NI = 32768, 512


def _say(msg):
    print("[gpu-multiset-case] " + msg, flush=True)


def _fields1d():
    """(beam, pixsize, [(data, sigma, binned, like_norm), ...]) per set."""
    gb, gps = synth.gaussian_beam()
    nb, nps = fitsio.read_beam(os.path.join(REFDATA,
                                            "band12_beam_f100.fits[0]"))
    maps = []
    for side, sig, seed in ((200, 0.0025, 7), (150, 0.0035, 8),
                            (160, 0.003, 9), (180, 0.002, 10)):
        d, _ = synth.make_map(side, sig, seed=seed)
        maps.append((d, sig))
    return [(gb.ravel(), gps, [(maps[0][0], maps[0][1], True, 1.0),
                               (maps[1][0], maps[1][1], False, 1.0)]),
            (nb, nps, [(maps[2][0], maps[2][1], False, 2.5),
                       (maps[3][0], maps[3][1], True, 1.0)])]


def test_two_beamsets_two_datasets_1d():
    sets = _fields1d()
    knots = synth.knot_positions()
    th = synth.walker_thetas(8, seed=21)
    ctx = _lib.Context(0)
    ctx.configure(NF, NI, knots, 0.0)
    L = O.Like(NF, NI, knots)
    keep = []
    for beam, ps, fields in sets:
        gbm = _lib.Beam(beam, ps, 1e-6, 120)
        obm = O.Beam(beam, ps, 1e-6, 120)
        keep += [gbm, obm]
        bg = ctx.add_beamset(gbm)
        bo = L.add_beamset(obm)
        for d, sig, binned, norm in fields:
            if binned:
                bins, c0, dl = _lib.data_bin(d, 1000)
                ctx.add_field(bg, bins=bins, bincent0=c0, bindelta=dl,
                              sigma_base=sig, like_norm=norm, npix=d.size)
                obins, oc0, odl = O.data_bin(d, 1000)
                assert np.array_equal(bins, obins) and c0 == oc0 and dl == odl
                L.add_field(bo, bins=obins, bincent0=oc0, bindelta=odl,
                            sigma_base=sig, like_norm=norm, npix=d.size)
            else:
                ctx.add_field(bg, pix=d, sigma_base=sig, like_norm=norm)
                L.add_field(bo, pix=d, sigma_base=sig, like_norm=norm)
    ctx.prepare()
    L.finalize(th[0])
    for b in range(2):
        gi, oi = ctx.beamset_info(b), L.beamset_info(b)
        for k in ("minRFlux", "maxRFlux", "maxsigma_base", "minDataFlux",
                  "maxDataFlux"):
            assert gi[k] == oi[k], (b, k)
    assert ctx.beamset_info(1)["minRFlux"] < 0  # the two-signed beam
    _say("1D: 2 beam sets x 2 datasets, 8 walkers")
    ll, st = ctx.loglike_batch(th)
    ref, rst = L.loglike_batch(th, 8)
    assert np.all(st == 0) and np.all(rst == 0), (st, rst)
    np.testing.assert_allclose(ll, ref, rtol=0, atol=1e-6)
    # the split transforms (several workgroups per walker) give the same
    # bits as one workgroup per walker
    for parts in (1, 2):
        ctx.set_parts(parts)
        llp, stp = ctx.loglike_batch(th)
        assert np.all(stp == 0)
        np.testing.assert_array_equal(llp, ll)
    ctx.set_parts(0)


def test_two_beamsets_one_binned_dataset_each_1d():
    """A beam set with a single binned dataset does its dataset sum in the
    finalize kernel (no combine launch); the second beam set adds to the
    first's log likelihood.  Against the oracle at 1e-6, with a walker
    whose knot value is NaN (invalid model in both beam sets) and one with
    a negative sigma multiplier (bad_like, 1e25 per beam set,
    src/calcLike.cc:403), and the same bits for every walker split."""
    gb, gps = synth.gaussian_beam()
    nb, nps = fitsio.read_beam(os.path.join(REFDATA,
                                            "band12_beam_f100.fits[0]"))
    d0, _ = synth.make_map(200, 0.0025, seed=11)
    d1, _ = synth.make_map(160, 0.003, seed=12)
    knots = synth.knot_positions()
    th = synth.walker_thetas(8, seed=23)
    th[3, 2] = np.nan
    th[5, 8] = -1.0
    ctx = _lib.Context(0)
    ctx.configure(NF, NI, knots, 0.0)
    L = O.Like(NF, NI, knots)
    keep = []
    for beam, ps, d, sig, norm in ((gb.ravel(), gps, d0, 0.0025, 1.0),
                                   (nb, nps, d1, 0.003, 2.0)):
        gbm = _lib.Beam(beam, ps, 1e-6, 120)
        obm = O.Beam(beam, ps, 1e-6, 120)
        keep += [gbm, obm]
        bg = ctx.add_beamset(gbm)
        bo = L.add_beamset(obm)
        bins, c0, dl = _lib.data_bin(d, 1000)
        ctx.add_field(bg, bins=bins, bincent0=c0, bindelta=dl,
                      sigma_base=sig, like_norm=norm, npix=d.size)
        L.add_field(bo, bins=bins, bincent0=c0, bindelta=dl,
                    sigma_base=sig, like_norm=norm, npix=d.size)
    ctx.prepare()
    L.finalize(th[0])
    _say("1D: 2 beam sets x 1 binned dataset, 8 walkers")
    ll, st = ctx.loglike_batch(th)
    ref, rst = L.loglike_batch(th, 8)
    ok = [w for w in range(8) if w not in (3, 5)]
    assert np.all(st[ok] == 0) and np.all(rst[ok] == 0), (st, rst)
    np.testing.assert_allclose(ll[ok], ref[ok], rtol=0, atol=1e-6)
    assert st[3] == 2 and rst[3] == 2 and np.isnan(ll[3])  # invalid model
    assert st[5] == 4 and ll[5] == 2e25 and ref[5] == 2e25  # bad_like x 2
    for parts in (1, 2, 4):
        ctx.set_parts(parts)
        llp, stp = ctx.loglike_batch(th)
        np.testing.assert_array_equal(stp, st)
        np.testing.assert_array_equal(llp[ok], ll[ok])
    ctx.set_parts(0)


def test_two_beamsets_two_datasets_2d():
    n, nedge = 512, 64
    th0 = synth2d.true_theta()
    th = synth2d.walker_thetas(4, seed=23)
    kp, sp, op = synth2d.positions()
    ctx = _lib.Context2(0)
    ctx.configure2d(n, kp, sp, op, nedge=nedge, edge_integ=True)
    L = O2.Like2(n, nedge, True)
    L.set_positions(kp, sp, op)
    keep = []
    for b, (side, seed, nbhist, sig) in enumerate(
            ((90, 31, 120, (synth2d.SIGMA1, synth2d.SIGMA2)),
             (80, 32, 100, (0.0032, 0.0035)))):
        inp = make_inputs2d(npix_side=side, seed=seed, nbeamhist=nbhist,
                            nbins=50)
        gbm = _lib.DoubleBeam(inp["beam1"], inp["beam2"], inp["pixsize"],
                              inp["minbeamval"], inp["nbeamhist"])
        obm = O2.DoubleBeam(inp["beam1"], inp["beam2"], inp["pixsize"],
                            inp["minbeamval"], inp["nbeamhist"])
        keep += [gbm, obm, inp]
        bg = ctx.add_beamset2d(gbm)
        bo = L.add_beamset(obm)
        # dataset 0 unbinned, dataset 1 binned with another noise level
        ctx.add_field2d(bg, inp["d1"], inp["d2"], sig[0], sig[1])
        L.add_field(bo, inp["d1"], inp["d2"], sig[0], sig[1])
        s2 = (sig[0] * 1.3, sig[1] * 0.9)
        ctx.add_field2d(bg, inp["d1"], inp["d2"], s2[0], s2[1],
                        like_norm=1.5, bins=inp["bins"],
                        bininfo=inp["bininfo"])
        L.add_field(bo, inp["d1"], inp["d2"], s2[0], s2[1], like_norm=1.5,
                    bins=inp["bins"], bininfo=inp["bininfo"])
    ctx.prepare2d(th0)
    L.finalize(th0)
    _say("2D: 2 beam sets x 2 datasets, n=512, 4 walkers")
    ll, st = ctx.loglike2d_batch(th)
    ref, rst = L.loglike_batch(th, 4)
    assert np.all(st == 0) and np.all(rst == 0), (st, rst)
    np.testing.assert_allclose(ll, ref, rtol=0, atol=1e-6)
