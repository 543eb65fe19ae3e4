"""The CPU oracle against the reference's own golden vectors.

Each test reproduces an assertion of the reference test suite on the same
fixture files (copied verbatim into tests/golden/reference_testdata/) with the
reference's tolerance.  Sources: test/beam_test.cc, test/model_test.cc,
test/data_test.cc, test/pofd_test.cc.
"""
import math
import os

import numpy as np
import pytest

import pyoracle as O
from pofd_affine_amd import fitsio


def _beam(refdata, name, minval=1e-5, nhist=None):
    pix, ps = fitsio.read_beam(os.path.join(refdata, name))
    return O.Beam(pix, ps, minval, nhist)


# ---------------- test/beam_test.cc ----------------------------------------
def test_beam_read(refdata):
    # beam_test.cc:15-41
    b = _beam(refdata, "band1_beam.fits")
    i = b.info()
    assert i["haspos"] and not i["hasneg"]
    assert i["npos"] == 277 and i["nneg"] == 0
    assert not i["is_pos_hist"] and not i["is_neg_hist"]
    assert abs(i["pixsize"] - 4.0) < 1e-4
    assert abs(i["area"] - 2.8327039e-5) < 1e-8
    assert abs(i["area_pos"] - 2.8327039e-5) < 1e-8
    assert i["area_neg"] == 0.0
    assert abs(i["area_pix"] - 22.944899) < 1e-4
    assert abs(i["area_sq"] - 1.3852891e-5) < 1e-8
    assert abs(i["maxpos"] - 0.97798554) < 1e-5


def test_beam_histogram(refdata):
    # beam_test.cc:112-148
    b = _beam(refdata, "band1_beam.fits", 1e-5, 120)
    i = b.info()
    assert i["is_pos_hist"]
    assert i["posnbins"] == 39
    assert i["minval"] == 1e-5
    assert abs(i["area_pix"] - 22.944899) < 1e-6
    w = b.array("posweights")
    v = b.array("poshistval")
    np.testing.assert_array_equal(w[34:39], [8.0, 4.0, 4.0, 4.0, 1.0])
    np.testing.assert_allclose(v[34:39],
                               [1.99712, 1.74686, 1.33647, 1.169, 1.02251],
                               atol=1e-4, rtol=0)
    # weights count every kept pixel exactly once
    assert w.sum() == i["npos"]


def test_beam_negative(refdata):
    # beam_test.cc:151-172
    b = _beam(refdata, "band12_beam_f100.fits[0]", 1e-7)
    i = b.info()
    assert i["minval"] == 1e-7
    assert i["npos"] == 477 and i["nneg"] == 364
    assert abs(i["pixsize"] - 8.0) < 1e-4
    assert abs(i["area"] - 1.0765914e-4) < 1e-8
    assert abs(i["area_pos"] - 5.3829571e-5) < 1e-8
    assert abs(i["area_neg"] - 5.3829571e-5) < 1e-8
    assert abs(i["maxpos"] - 0.82171527) < 1e-5
    assert abs(b.array("negpix")[-1] - 0.082822058) < 1e-6


# ---------------- test/model_test.cc (1D part) ------------------------------
KP = [0.002, 0.005, 0.010, 0.020, 0.050]
KV = [10.0, 7.0, 5.0, 4.0, -1.0]


def test_model_number_counts():
    # model_test.cc:167-193 (knots as float in the reference test)
    kp = np.array(KP, dtype=np.float32).astype(np.float64)
    m = O.Knots(kp, KV)
    assert m.is_valid()
    for i in range(1, 4):
        assert abs(math.log10(m.number_counts(kp[i])[0]) - KV[i]) < 1e-4
    got = np.log10(m.number_counts([0.003, 0.007, 0.015]))
    np.testing.assert_allclose(got, [8.69422, 5.88217, 4.53303], atol=1e-3,
                               rtol=0)


def test_model_rrange(refdata):
    # model_test.cc:217-236
    kp = np.array(KP, dtype=np.float32).astype(np.float64)
    m = O.Knots(kp, KV)
    lo, hi = m.rrange(_beam(refdata, "band1_beam.fits"))
    assert lo == 0.0
    assert abs(hi - 0.0488993) < 1e-4


def test_model_integration():
    # model1DTest.Integration (model_test.cc:197-213)
    kp = np.array(KP, dtype=np.float32).astype(np.float64)
    m = O.Knots(kp, KV)
    assert abs(m.splineint(0) - 3.12022e6) < 1e4
    assert abs(m.splineint(1) - 7386.67) < 1.0
    assert abs(m.splineint(2) - 18.132) < 0.01


@pytest.mark.parametrize("nhist", [None, 120])
def test_model_getR(refdata, nhist):
    # model_test.cc:242-316
    kp = np.array(KP, dtype=np.float32).astype(np.float64)
    m = O.Knots(kp, KV)
    b = _beam(refdata, "band1_beam.fits", 1e-5, nhist)
    f = [0.003, 0.011, 0.03]
    want = np.array([1891.52, 0.817009, 0.000313631])
    for got in (m.getR1(f, b), m.getR(f, b)):
        assert np.all(np.abs(got - want) / want < 1e-3)
    if nhist is None:
        assert m.getR1([-0.001], b)[0] == 0.0
        assert m.getR1([0.1], b)[0] == 0.0


@pytest.mark.parametrize("nhist", [None, 200])
def test_model_getR_negative_beam(refdata, nhist):
    # model_test.cc:320-394
    kp = np.array(KP, dtype=np.float32).astype(np.float64)
    m = O.Knots(kp, KV)
    b = _beam(refdata, "band12_beam_f100.fits[0]", 1e-5, nhist)
    if nhist:
        assert b.info()["is_pos_hist"] and b.info()["is_neg_hist"]
    f = [-0.03, -0.003, -1e-4, 0.003, 0.03]
    want = np.array([0.0, 0.019075718, 11916638.0, 969.44687, 7.2612855e-5])
    for got in (m.getR1(f, b), m.getR(f, b)):
        rel = np.where(want > 0, np.abs(got - want) / np.where(want > 0, want,
                                                                1.0),
                       np.abs(got - want))
        assert np.all(rel < 1e-3), (got, want)


# ---------------- test/data_test.cc ------------------------------------------
def test_data_read(refdata):
    # data_test.cc:10-47 (data reads are host ingestion, fitsio.read_map)
    p = os.path.join(refdata, "testmodel2D_band1.fits")
    d = fitsio.read_map(p)
    assert d.size == 38431
    assert abs(d.min() + 0.011066) < 1e-5 and abs(d.max() - 0.0659472) < 1e-5
    d = fitsio.read_map(p, ignore_mask=True)
    assert d.size == 40401
    assert abs(d.min() + 0.011066) < 1e-5 and abs(d.max() - 0.0659472) < 1e-5


def test_data_read_smap(refdata):
    # data_test.cc:50-87
    p = os.path.join(refdata, "test_PSW.fits")
    d = fitsio.read_map(p, ignore_mask=True)
    assert d.size == 8100
    assert abs(d.min() + 0.0063209710) < 1e-5
    assert abs(d.max() - 0.052431226) < 1e-5
    d = fitsio.read_map(p)
    assert d.size == 5355
    assert abs(d.min() + 0.0063209710) < 1e-5
    assert abs(d.max() - 0.046583797) < 1e-5


# ---------------- test/pofd_test.cc (1D part) --------------------------------
def test_pd_gaussian_fill():
    # pofd_test.cc:20-80
    n = 1024
    i = np.arange(n, dtype=np.float64)
    dat = np.exp(-(i - 200.0) * (i - 200.0) / 5000.0)
    minflux, dflux = -0.3, 1e-3
    for idx in (0, 20, 133, 1023):
        assert O.pd_interp(dat, minflux, dflux,
                           minflux + idx * dflux)[0] == pytest.approx(
                               dat[idx], rel=1e-6)
    for idx in (0, 20, 133):
        x = minflux + idx * dflux + 0.5 * dflux
        assert O.pd_interp(dat, minflux, dflux, x)[0] == pytest.approx(
            0.5 * (dat[idx] + dat[idx + 1]), rel=1e-6)
    mn = O.pd_mean(dat, minflux, dflux, False, True)
    assert abs(mn - (minflux + 200 * dflux)) < 1e-4
    m, v = O.pd_mean_var(dat, minflux, dflux, False, True)
    assert abs(m - (minflux + 200 * dflux)) < 1e-4
    assert abs(v - 2.5e-3) < 1e-5
    lg = np.where(dat > 0, np.log2(np.where(dat > 0, dat, 1.0)), -30.0)
    for idx in (160, 200, 230, 255):
        assert lg[idx] == pytest.approx(math.log2(dat[idx]), rel=1e-7)
    mn = O.pd_mean(lg, minflux, dflux, True, True)
    assert abs(mn - (minflux + 200 * dflux)) < 1e-4
    m, v = O.pd_mean_var(lg, minflux, dflux, True, True)
    assert abs(m - (minflux + 200 * dflux)) < 1e-4
    assert abs(v - 2.5e-3) < 1e-5
