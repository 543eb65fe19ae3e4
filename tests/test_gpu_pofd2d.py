"""GPU parity of the 2D (two-band) P(D) likelihood against the 2D oracle.

Every call goes through the C ABI (include/pofd_affine_amd.h, pofd2d_*).
logL within 1e-6 absolute where the double-precision transforms allow it;
the P(D) within 1e-10 of its peak; R moments, split points and the
normalisation agree to rounding.
"""
import numpy as np
import pytest

from pofd_affine_amd import _lib, synth2d, workload2d
from oracle2d_side import build_oracle2d, make_inputs2d

pytestmark = pytest.mark.gpu

N_S, NEDGE_S = 512, 64


def _say(msg):
    print("[gpu2d-case] " + msg, flush=True)


@pytest.fixture(scope="module")
def small2d():
    inp = make_inputs2d(npix_side=120, seed=5, nbins=60)
    th0 = synth2d.true_theta()
    th = synth2d.walker_thetas(6, seed=17)
    out = {}
    for binned in (False, True):
        _say("build small 2D binned=%s" % binned)
        ctx = workload2d.build_context2d(inp, N_S, th0, nedge=NEDGE_S,
                                         binned=binned)
        L = build_oracle2d(inp, N_S, th0, nedge=NEDGE_S, binned=binned)
        ll, st = ctx.loglike2d_batch(th)
        out[binned] = (ctx, L, ll, st)
    return inp, th, out


def test_small2d_setup_matches(small2d):
    inp, th, out = small2d
    ctx, L, _, _ = out[False]
    g = ctx.beamset_info2d(0)
    o = L.beamset_info(0)
    for k in ("minR1", "maxR1", "minR2", "maxR2", "maxsig1", "maxsig2",
              "mind1", "maxd1", "mind2", "maxd2"):
        assert g[k] == o[k], k
    assert g["like_offset0"] == o["like_offset0"]


@pytest.mark.parametrize("binned", [False, True])
def test_small2d_loglike(small2d, binned):
    inp, th, out = small2d
    ctx, L, ll, st = out[binned]
    _say("small 2D loglike binned=%s" % binned)
    ref, rst = L.loglike_batch(th, 6)
    assert np.all(rst == 0) and np.all(st == 0), (st, rst)
    np.testing.assert_allclose(ll, ref, rtol=0, atol=1e-6)


@pytest.mark.parametrize("w", [0, 4])
def test_small2d_moments_and_pd(small2d, w):
    inp, th, out = small2d
    ctx, L, _, _ = out[False]
    _say("small 2D P(D) walker %d" % w)
    ip = L.initpd(0, th[w], want_arrays=False)
    pd_ref, info_ref = L.getpd(0, 0, float(th[w][-6]), float(th[w][-5]),
                               setlog=False)
    dbg = ctx.get_debug2d(w)
    for a, b in (("mn1", "mn1"), ("mn2", "mn2"), ("var1", "var1"),
                 ("var2", "var2")):
        assert dbg[a] == pytest.approx(ip[b], rel=1e-12), a
    pd, info = ctx.get_pd2d(w, is_log=False)
    assert info["split1"] == info_ref["split1"]
    assert info["split2"] == info_ref["split2"]
    assert info["minflux1"] == pytest.approx(info_ref["minflux1"], abs=1e-12)
    assert info["minflux2"] == pytest.approx(info_ref["minflux2"], abs=1e-12)
    np.testing.assert_allclose(pd, pd_ref, rtol=0, atol=1e-10 * pd_ref.max())


def test_small2d_bonus_integrals(small2d):
    """getNS, getFluxPerArea and getFluxSqPerArea of both bands per walker
    (numberCountsDoubleLogNormal::splineInt) against the oracle."""
    import pyoracle2d as O2
    inp, th, out = small2d
    ctx = out[False][0]
    _say("small 2D bonus integrals")
    bo = ctx.get_bonus2d(len(th))
    kp, sp, op = synth2d.positions()
    nmp = len(kp) + len(sp) + len(op)
    for w in range(len(th)):
        m = O2.LogNormalModel(kp, sp, op)
        m.set_params(th[w][:nmp])
        ref = [m.splineint(a, b) for a, b in ((0, 0), (1, 0), (0, 1),
                                                (2, 0), (0, 2))]
        np.testing.assert_allclose(bo[w], ref, rtol=1e-12, atol=0)


def test_small2d_status_codes(small2d):
    inp, th, out = small2d
    ctx = out[False][0]
    bad = th[:3].copy()
    nmp = len(synth2d.KNOTS_MJY) + len(synth2d.SIGMA_VALS) + len(
        synth2d.OFFSET_VALS)
    bad[0, 1] = np.nan             # invalid model
    bad[1, nmp] = -1.0             # negative sigma multiplier: bad_like
    ll, st = ctx.loglike2d_batch(bad)
    assert st[0] == _lib.POFD_E_INVALID_MODEL and np.isnan(ll[0])
    assert st[1] == _lib.POFD_E_BADLIKE and ll[1] == 1e25
    assert st[2] == _lib.POFD_OK


def test_small2d_deterministic(small2d):
    inp, th, out = small2d
    ctx, _, ll, _ = out[False]
    ll2, _ = ctx.loglike2d_batch(th)
    assert np.array_equal(ll, ll2)


def _full2d(binned, nwalk=4):
    """The measured 2D configuration: n = 2048 grid, 1000x1000 pixel pairs,
    120x120-bin beam histogram, sigma 2.5 / 2.7 mJy."""
    inp = make_inputs2d(npix_side=1000, seed=99, nbeamhist=120,
                        minbeamval=1e-6, nbins=500 if binned else None)
    th0 = synth2d.true_theta()
    th = synth2d.walker_thetas(nwalk, seed=4321)
    ctx = workload2d.build_context2d(inp, 2048, th0, nedge=256,
                                     binned=binned)
    ll, st = ctx.loglike2d_batch(th)
    L = build_oracle2d(inp, 2048, th0, nedge=256, binned=binned)
    ref, rst = L.loglike_batch(th, nwalk)
    return ll, st, ref, rst


@pytest.mark.parametrize("binned", [False, True])
def test_config2d_full_size(binned):
    _say("full size 2D n=2048 binned=%s" % binned)
    ll, st, ref, rst = _full2d(binned)
    assert np.all(st == 0) and np.all(rst == 0), (st, rst)
    print("[gpu2d-case] max |dlogL| = %.3e" % np.max(np.abs(ll - ref)),
          flush=True)
    np.testing.assert_allclose(ll, ref, rtol=0, atol=1e-6)
