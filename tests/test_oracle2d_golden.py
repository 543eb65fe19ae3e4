"""Pins of the 2D oracle (oracle/pofd_oracle2d.c) against the golden values
of the reference's own tests (test/beam_test.cc beam2DTest,
test/model_test.cc model2DTest, test/pofd_test.cc pd2DTest,
test/data_test.cc fitsDataDoubleTest), read from the reference's FITS
fixtures kept in tests/golden/reference_testdata/."""
import math
import os

import numpy as np
import pytest

import pyoracle as O
import pyoracle2d as O2
from pofd_affine_amd import fitsio


def _beam2(refdata, f1, f2, minval, nhist=None):
    b1, ps1 = fitsio.read_beam(os.path.join(refdata, f1))
    b2, ps2 = fitsio.read_beam(os.path.join(refdata, f2))
    assert abs((ps1 - ps2) / ps1) <= 1e-5
    return O2.DoubleBeam(b1, b2, ps1, minval, nhist)


def _area(tot, pixsize):
    return tot * (pixsize / 3600.0) ** 2


def test_beam2d_read(refdata):
    # beam2DTest.Read (test/beam_test.cc:177-245)
    bm = _beam2(refdata, "band1_beam.fits", "band2_beam.fits", 0.0)
    s, g = bm.info()
    assert [x["hassign"] for x in s] == [1, 0, 0, 0]
    assert s[0]["npix"] == 625
    assert abs(g["pixsize"] - 4.0) < 1e-4
    bm = _beam2(refdata, "band1_beam.fits", "band2_beam.fits", 1e-6)
    s, g = bm.info()
    assert s[0]["npix"] == 325 and sum(x["npix"] for x in s) == 325
    assert [x["ishist"] for x in s] == [0, 0, 0, 0]
    assert abs(_area(s[0]["tot1"], g["pixsize"]) - 2.8327225e-5) < 1e-8
    assert abs(s[0]["max1"] - 0.97798554) < 1e-5
    assert abs(_area(s[0]["tot2"], g["pixsize"]) - 5.4604548e-5) < 1e-8
    assert abs(s[0]["max2"] - 0.98850347) < 1e-5


def test_beam2d_histogram(refdata):
    # beam2DTest.Histogram (:247-264)
    bm = _beam2(refdata, "band1_beam.fits", "band2_beam.fits", 1e-6, 150)
    s, g = bm.info()
    assert g["nbins"] == 150
    assert [x["ishist"] for x in s] == [1, 0, 0, 0]
    assert [x["nhist"] for x in s] == [45, 0, 0, 0]
    # weights are pixel counts; 1/<beam> per bin lies inside the bin's range
    w = bm.array(0, "binw")
    assert w.sum() == 325
    v1 = 1.0 / bm.array(0, "binv1")
    assert np.all(v1 >= s[0]["min1"] * (1 - 1e-12))
    assert np.all(v1 <= s[0]["max1"] * (1 + 1e-12))


def test_beam2d_negative(refdata):
    # beam2DTest.NegBeam (:266-326)
    bm = _beam2(refdata, "band12_beam_f100.fits[0]",
                "band12_beam_f100.fits[1]", 1e-7)
    s, g = bm.info()
    assert [x["hassign"] for x in s] == [1, 1, 1, 1]
    assert [x["npix"] for x in s] == [469, 8, 16, 348]
    assert abs(g["pixsize"] - 8.0) < 1e-4
    areas = [5.3796224e-5, 3.3357658e-8, 3.3918851e-6, 5.0437686e-5]
    for k in range(4):
        assert abs(_area(s[k]["tot1"], g["pixsize"]) - areas[k]) < 1e-8
    assert abs(s[0]["min1"] - 0.0010022319) < 1e-6
    assert abs(s[0]["max1"] - 0.82171527) < 1e-6
    assert abs(s[0]["min2"] - 0.0022131729) < 1e-6
    assert abs(s[0]["max2"] - 0.74992756) < 1e-6
    assert abs(s[2]["min1"] - 0.019993591) < 1e-6
    assert abs(s[2]["max1"] - 0.055385745) < 1e-6
    assert abs(s[2]["min2"] - 0.0015057219) < 1e-6
    assert abs(s[2]["max2"] - 0.065584647) < 1e-6
    bm2 = _beam2(refdata, "band12_beam_f100.fits[0]",
                 "band12_beam_f100.fits[1]", 1e-7, 200)
    s, _ = bm2.info()
    assert [x["ishist"] for x in s] == [1, 0, 0, 1]


KP = [0.002, 0.005, 0.010, 0.020, 0.040]


def _model(sigmapos, offsetpos, pvals):
    m = O2.LogNormalModel(np.float32(KP).astype(np.float64),
                          np.float32(sigmapos).astype(np.float64),
                          np.float32(offsetpos).astype(np.float64))
    m.set_params(np.float32(pvals))
    return m


def test_model2d_params_and_range(refdata):
    # model2DTest.SetGetParams / getModelRange / getRRange (:535-646)
    m = _model([0.002, 0.020], [0.03],
               [8.0, 6.0, 4.0, 3.3, 2.0, 0.1, 0.2, -0.03])
    assert m.is_valid()
    mx = m.maxflux()
    assert abs(mx[0] - 0.040) < 1e-5
    assert abs(mx[1] - 0.067957) < 1e-3
    bm = _beam2(refdata, "band1_beam.fits", "band2_beam.fits", 1e-6)
    r = m.rrange(bm)
    assert r[0] == 0 and r[2] == 0
    assert abs(r[1] - 0.039119421) < 1e-4
    assert abs(r[3] - 0.0671758809) < 1e-4
    # invalid: NaN knot value
    m2 = _model([0.002, 0.020], [0.03],
                [8.0, float("nan"), 4.0, 3.3, 2.0, 0.1, 0.2, -0.03])
    assert not m2.is_valid()


def test_model2d_number_counts():
    # model2DTest.NumberCounts (:648-679)
    m = _model([0.002, 0.020], [0.03],
               [8.0, 6.0, 4.0, 3.3, 2.0, 0.1, 0.2, -0.03])
    f1 = [0.003, 0.011, 0.011, 0.03]
    f2 = [0.003, 0.010, 0.015, 0.028]
    exp = [2.09837e+10, 1.67999e+06, 94125.8, 29739.2]
    for a, b, e in zip(f1, f2, exp):
        assert abs(m.counts(a, b) - e) / e < 1e-3


def test_model2d_power_integration():
    # model2DTest.PowerIntegration (:681-703)
    m = _model([0.002, 0.020], [0.03],
               [8.0, 6.0, 4.0, 3.3, 2.0, 0.1, 0.2, -0.03])
    assert abs(m.splineint(0, 0) - 57266.8) < 10.0
    assert abs(m.splineint(1, 0) - 152.262) < 0.1
    assert abs(m.splineint(0, 1) - 148.585) < 0.1


def test_model2d_band1_projection_matches_1d_spline():
    # model2DTest.Projection band 1 (:706-744): the band-1 counts are the
    # 1D knot spline's
    pv = [8.0, 6.0, 4.0, 3.3, 2.0, 0.1, -0.03]
    m = _model([0.003], [0.03], pv)
    k1 = O.Knots(np.float32(KP).astype(np.float64))
    k1.set_params(np.float32(pv[:5]))
    for i in range(1, 4):
        assert abs(pv[i] - math.log10(m.band1_counts(KP[i]))) < 1e-4
    grid = KP[0] + (KP[-1] - KP[0]) / 99.0 * np.arange(1, 99)
    for f in grid:
        assert math.log10(m.band1_counts(f)) == pytest.approx(
            math.log10(float(np.ravel(k1.number_counts(f))[0])), abs=1e-12)


def _check_R(m, bm, f1, f2, exp, f1v, f2v, exp2d):
    for a, b, e in zip(f1, f2, exp):
        assert abs(m.getR1(a, b, bm) - e) / e < 1e-3, (a, b)
    R = m.getR(f1v, f2v, bm)
    for i in range(len(f1v)):
        for j in range(len(f2v)):
            e = exp2d[i * len(f2v) + j]
            d = abs(R[i, j] - e) / e if e > 0 else abs(R[i, j] - e)
            assert d < 1e-3, (i, j, R[i, j], e)


def test_model2d_getR(refdata):
    # model2DTest.getR (:761-822)
    m = _model([0.015], [0.03], [8.0, 6.0, 4.0, 3.3, 2.0, 0.15, -0.03])
    bm = _beam2(refdata, "band1_beam.fits", "band2_beam.fits", 1e-6)
    _check_R(m, bm, [0.003, 0.011, 0.011, 0.03], [0.003, 0.010, 0.015, 0.028],
             [76926.027, 9.6701326, 11.520839, 0.12455423],
             [0.003, 0.011, 0.02], [0.003, 0.015],
             [76926.027, 0.024892799, 1.2063052e-15, 11.520839,
              1.9838477e-34, 0.21958590])


def test_model2d_getR_hist(refdata):
    # model2DTest.getRHist (:824-887), 150-bin histogram
    m = _model([0.020], [0.03], [8.0, 6.0, 4.0, 3.3, 2.0, 0.15, -0.03])
    bm = _beam2(refdata, "band1_beam.fits", "band2_beam.fits", 1e-6, 150)
    _check_R(m, bm, [0.003, 0.011, 0.011, 0.03], [0.003, 0.010, 0.015, 0.028],
             [76926.027, 9.6701326, 11.520839, 0.12455423],
             [0.003, 0.011, 0.02], [0.003, 0.015],
             [76926.027, 0.024892799, 1.2063052e-15, 11.520839,
              1.9838477e-34, 0.21958590])


@pytest.mark.parametrize("nhist", [None, 200])
def test_model2d_getR_negative_beam(refdata, nhist):
    # model2DTest.getRNeg (:889-995)
    m = _model([0.015], [0.03], [8.0, 6.0, 4.0, 3.3, 2.0, 0.15, -0.3])
    bm = _beam2(refdata, "band12_beam_f100.fits[0]",
                "band12_beam_f100.fits[1]", 1e-7, nhist)
    _check_R(m, bm, [-2e-4, -2e-4, 0.003, 0.011], [-3e-4, 1e-4, 0.002, 0.010],
             [1.596908e8, 12727124, 75840.624, 20.409118],
             [-3e-4, -1e-4, 1e-3], [-2e-4, 0.5e-3],
             [37281355, 771.54407, 45388812, 49.613048, 0.0, 7762.3524])
    for a, b in [(0.05, 0.03), (0.05, -0.03), (-0.05, 0.03), (-0.05, -0.03)]:
        assert m.getR1(a, b, bm) == 0.0


def test_pd2d_means():
    # pd2DTest.Data (test/pofd_test.cc:94-140)
    n = 512
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    dat = np.exp(-((i - 180.0) ** 2 + (j - 223.0) ** 2) / 800.0)
    mn = O2.pd2_means(dat, -0.3, 1e-3, -0.1, 3e-3, donorm=True)
    assert abs(mn[0] - (-0.3 + 1e-3 * 180)) < 1e-4
    assert abs(mn[1] - (-0.1 + 3e-3 * 223)) < 1e-4


def test_data2d_read(refdata):
    # fitsDataDoubleTest.Read (test/data_test.cc:83-130)
    f1 = os.path.join(refdata, "testmodel2D_band1.fits")
    f2 = os.path.join(refdata, "testmodel2D_band2.fits")
    d1, d2 = fitsio.read_map_pair(f1, f2)
    assert d1.size == 38431
    assert abs(d1.min() - -0.011066) < 1e-5
    assert abs(d2.min() - -0.0176356) < 1e-5
    assert abs(d1.max() - 0.0659472) < 1e-5
    assert abs(d2.max() - 0.0776496) < 1e-5
    d1, d2 = fitsio.read_map_pair(f1, f2, ignore_mask=True)
    assert d1.size == 40401
