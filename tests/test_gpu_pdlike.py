"""PD::getLogLike and PDDouble::getLogLike on the device (pofd_pd_loglike,
pofd_pd2_loglike) against the oracle's restatement of src/PD.cc:538-677
and src/PDDouble.cc:976-1222, which is pinned by closed forms
(test_oracle_pdlike.py).  The P(D)s are real ones from the oracle's
pipeline, stored as log2 and linearly; the data include points off both
ends of the grid.  1e-6 absolute (the sums differ only in order)."""
import numpy as np
import pytest

import pyoracle as O
import pyoracle2d as O2
from oracle_side import build_oracle
from oracle2d_side import build_oracle2d, make_inputs2d
from pofd_affine_amd import _lib, synth, synth2d, workload

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-pdlike-case] " + msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    return _lib.Context(0)


@pytest.mark.parametrize("islog", [True, False])
def test_pd_loglike_1d(ctx, islog):
    inp = workload.make_inputs(npix_side=200, seed=7)
    th = synth.walker_thetas(2, seed=11)
    L = build_oracle(inp, 32768, 512, True, th[0])
    L.initpd(0, th[1])
    r = L.getpd(0, 0, float(th[1][8]), setlog=islog)
    pd, mf, df = r["pd"], r["minflux"], r["dflux"]
    d = np.concatenate([inp["data"], [mf - 0.01, mf, mf + (pd.size - 1.5) * df,
                                      mf + pd.size * df + 0.02]])
    _say("1D unbinned, islog=%s" % islog)
    want, _ = O.pd_loglike_unbinned(pd, mf, df, d, islog=islog)
    got = ctx.pd_loglike(pd, mf, df, is_log=islog, pix=d)
    assert got == pytest.approx(want, rel=0, abs=1e-6)
    bins, c0, dl = O.data_bin(d, 1000)
    _say("1D binned, islog=%s" % islog)
    want = O.pd_loglike_binned(pd, mf, df, bins, c0, dl, islog=islog)
    got = ctx.pd_loglike(pd, mf, df, is_log=islog, bins=bins, bincent0=c0,
                         bindelta=dl)
    assert got == pytest.approx(want, rel=0, abs=1e-6)


def test_pd_loglike_errors(ctx):
    with pytest.raises(_lib.PofdError, match="No data present"):
        ctx.pd_loglike(np.zeros(16), 0.0, 1.0, pix=np.zeros(0))


@pytest.mark.parametrize("islog", [True, False])
def test_pd_loglike_2d(ctx, islog):
    inp = make_inputs2d(npix_side=90, seed=5, nbins=40)
    th0 = synth2d.true_theta()
    L = build_oracle2d(inp, 512, th0, nedge=64)
    L.initpd(0, th0, want_arrays=False)
    pd, info = L.getpd(0, 0, float(th0[-6]), float(th0[-5]), setlog=islog)
    m1, m2 = info["minflux1"], info["minflux2"]
    d1, d2 = info["dflux1"], info["dflux2"]
    n1, n2 = pd.shape
    e1 = np.array([m1 - 0.01, m1 + (n1 - 1) * d1 + 0.01, 0.0, m1 - 1.0])
    e2 = np.array([0.0, m2 - 0.02, m2 + (n2 - 1) * d2 + 0.02, m2 + n2 * d2])
    x1 = np.concatenate([inp["d1"], e1])
    x2 = np.concatenate([inp["d2"], e2])
    _say("2D unbinned, islog=%s" % islog)
    want = O2.lib().orc_pd2_loglike_unbinned(
        n1, n2, O2._dp(O2._f64(pd)), m1, d1, m2, d2, int(islog),
        O2._dp(O2._f64(x1)), O2._dp(O2._f64(x2)), x1.size)
    got = ctx.pd2_loglike(pd, m1, d1, m2, d2, is_log=islog, d1=x1, d2=x2)
    assert got == pytest.approx(want, rel=0, abs=1e-6)
    bins, bi = O2.data2_bin(x1, x2, 40, 50)
    b = np.ascontiguousarray(bins, dtype=np.uint32)
    _say("2D binned, islog=%s" % islog)
    want = O2.lib().orc_pd2_loglike_binned(
        n1, n2, O2._dp(O2._f64(pd)), m1, d1, m2, d2, int(islog),
        b.ctypes.data_as(O2.U32P), 40, 50, *[float(v) for v in bi[:4]])
    got = ctx.pd2_loglike(pd, m1, d1, m2, d2, is_log=islog, bins=b,
                          bininfo=bi[:4])
    assert got == pytest.approx(want, rel=0, abs=1e-6)
