"""CPU-side checks of the C ABI library: it loads without a GPU, exports
every symbol the public header declares, and its host-side preparation
(beam histogram, data binning) is bit-exact with the oracle."""
import ctypes
import os

import numpy as np
import pytest

import pyoracle as O
from pofd_affine_amd import _lib, fitsio, synth


def test_library_exports_header_symbols():
    L = _lib.lib()
    names = _lib.exported_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n


def test_status_codes_match_oracle():
    assert (_lib.POFD_E_SPLIT, _lib.POFD_E_INVALID_MODEL, _lib.POFD_E_MEAN,
            _lib.POFD_E_BADLIKE, _lib.POFD_E_ARG) == (1, 2, 3, 4, 5)


def test_no_gpu_here_is_reported():
    # without a device the context cannot be created: the product path has
    # no CPU fallback
    try:
        import torch
        has = torch.cuda.is_available()
    except Exception:
        has = False
    if has:
        return
    p = ctypes.c_void_p()
    rc = _lib.lib().pofd_create(ctypes.byref(p), 0)
    assert rc == _lib.POFD_E_NOGPU


def _beams(refdata):
    out = []
    for name, mv in (("band1_beam.fits", 1e-5), ("band2_beam.fits", 1e-5),
                     ("band12_beam_f100.fits[0]", 1e-7)):
        pix, ps = fitsio.read_beam(os.path.join(refdata, name))
        out.append((pix, ps, mv))
    b2, ps = synth.gaussian_beam()
    out.append((b2.ravel(), ps, 1e-6))
    return out


def test_beam_prep_bit_exact(refdata):
    for pix, ps, mv in _beams(refdata):
        for nh in (None, 120, 200):
            g = _lib.Beam(pix, ps, mv, nh)
            o = O.Beam(pix, ps, mv, nh)
            gi, oi = g.info(), o.info()
            for k in gi:
                if not (np.isnan(gi[k]) and np.isnan(oi[k])):
                    assert gi[k] == oi[k], k
            for which in ("pospix", "negpix", "posinv", "neginv",
                          "posweights", "poshistval", "negweights",
                          "neghistval"):
                assert np.array_equal(g.array(which), o.array(which)), which


def test_data_binning_bit_exact(refdata):
    d = fitsio.read_map(os.path.join(refdata, "testmodel2D_band1.fits"))
    for data in (d, synth.make_map(300, seed=3)[0]):
        for nb in (1, 7, 1000):
            a = _lib.data_bin(data, nb)
            b = O.data_bin(data, nb)
            assert np.array_equal(a[0], b[0])
            assert a[1] == b[1] and a[2] == b[2]
    m1 = _lib.data_mean_subtract(d)
    m2 = O.data_mean_subtract(d)
    assert np.array_equal(m1[0], m2[0]) and m1[1] == m2[1]


def _fits_beam_c(path):
    L = _lib.lib()
    L.pofd_fits_read_beam.argtypes = [ctypes.c_char_p, ctypes.c_void_p,
                                      ctypes.c_uint64,
                                      ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_double)]
    n = ctypes.c_uint64()
    ps = ctypes.c_double()
    assert L.pofd_fits_read_beam(path.encode(), None, 0, ctypes.byref(n),
                                 ctypes.byref(ps)) == 0
    out = np.zeros(n.value)
    assert L.pofd_fits_read_beam(path.encode(), out.ctypes.data, n.value,
                                 ctypes.byref(n), ctypes.byref(ps)) == 0
    return out, ps.value


def _fits_map_c(path, ignore):
    L = _lib.lib()
    L.pofd_fits_read_map.argtypes = [ctypes.c_char_p, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_uint64,
                                     ctypes.POINTER(ctypes.c_uint64)]
    n = ctypes.c_uint64()
    assert L.pofd_fits_read_map(path.encode(), ignore, None, 0,
                                ctypes.byref(n)) == 0
    out = np.zeros(n.value)
    assert L.pofd_fits_read_map(path.encode(), ignore, out.ctypes.data,
                                n.value, ctypes.byref(n)) == 0
    return out


def test_fits_reader_c_matches_python(refdata, tmp_path):
    for name in ("band1_beam.fits", "band2_beam.fits",
                 "band12_beam_f100.fits[0]", "band12_beam_f100.fits[1]"):
        p = os.path.join(refdata, name)
        a, pa = _fits_beam_c(p)
        b, pb = fitsio.read_beam(p)
        assert np.array_equal(a, b) and pa == pb
    for name in ("testmodel2D_band1.fits", "test_PSW.fits"):
        p = os.path.join(refdata, name)
        for ig in (0, 1):
            assert np.array_equal(_fits_map_c(p, ig),
                                  fitsio.read_map(p, bool(ig)))
    # round trip of the writer, with a mask extension
    img = np.random.default_rng(1).normal(size=(13, 17))
    mask = (np.arange(13 * 17).reshape(13, 17) % 5 == 0).astype(np.int16)
    f = str(tmp_path / "m.fits")
    fitsio.write_image(f, img, extname="image", mask=mask)
    assert np.array_equal(fitsio.read_map(f), img.ravel()[mask.ravel() == 0])
    assert np.array_equal(_fits_map_c(f, 0), img.ravel()[mask.ravel() == 0])
    assert np.array_equal(_fits_map_c(f, 1), img.ravel())


def test_fits_map_pair_c_matches_python(refdata, tmp_path):
    """fitsDataDouble::readData's combined masks: the C reader and the
    Python one agree on the reference's two-band maps and on synthetic
    maps with one and with two (bitwise-combined) masks."""
    L = _lib.lib()
    L.pofd_fits_read_map_pair.argtypes = [
        ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p,
        ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]

    def pair_c(p1, p2, ig):
        n = ctypes.c_uint64(0)
        assert L.pofd_fits_read_map_pair(p1.encode(), p2.encode(), ig, None,
                                         None, 0, ctypes.byref(n)) == 0
        a, b = np.zeros(n.value), np.zeros(n.value)
        assert L.pofd_fits_read_map_pair(p1.encode(), p2.encode(), ig,
                                         a.ctypes.data, b.ctypes.data,
                                         n.value, ctypes.byref(n)) == 0
        return a, b

    p1 = os.path.join(refdata, "testmodel2D_band1.fits")
    p2 = os.path.join(refdata, "testmodel2D_band2.fits")
    for ig in (0, 1):
        a, b = pair_c(p1, p2, ig)
        ra, rb = fitsio.read_map_pair(p1, p2, bool(ig))
        assert np.array_equal(a, ra) and np.array_equal(b, rb)
    rng = np.random.default_rng(4)
    i1, i2 = rng.normal(size=(2, 11, 9))
    m1 = rng.integers(0, 4, size=(11, 9)).astype(np.int16)
    m2 = rng.integers(0, 4, size=(11, 9)).astype(np.int16)
    f1, f2, f3 = (str(tmp_path / x) for x in ("a.fits", "b.fits", "c.fits"))
    fitsio.write_image(f1, i1, extname="image", mask=m1)
    fitsio.write_image(f2, i2, extname="image", mask=m2)
    fitsio.write_image(f3, i2, extname="image")
    for q1, q2, keep in ((f1, f2, (m1 & m2) == 0), (f1, f3, m1 == 0),
                         (f3, f2, m2 == 0)):
        a, b = pair_c(q1, q2, 0)
        ra, rb = fitsio.read_map_pair(q1, q2)
        assert np.array_equal(a, ra) and np.array_equal(b, rb)
        assert a.size == int(keep.sum())


def test_dbeam_prep_bit_exact(refdata):
    """doublebeam::setBeams + makeHistogram in the library equal the oracle's
    restatement bit for bit, on the reference's two-band beam fixtures
    (including the four-sign-component beam)."""
    import pyoracle2d as O2
    for f1, f2, minval, nh in [("band1_beam.fits", "band2_beam.fits", 1e-6, 150),
                               ("band12_beam_f100.fits[0]",
                                "band12_beam_f100.fits[1]", 1e-7, 200)]:
        b1, ps = fitsio.read_beam(os.path.join(refdata, f1))
        b2, _ = fitsio.read_beam(os.path.join(refdata, f2))
        g = _lib.DoubleBeam(b1, b2, ps, minval, nh)
        o = O2.DoubleBeam(b1, b2, ps, minval, nh)
        gs, gg = g.info()
        os_, og = o.info()
        for a, b in zip(gs, os_):
            for k in a:
                assert (a[k] == b[k]) or (np.isnan(a[k]) and np.isnan(b[k])), k
        assert gg == og
        for s in range(4):
            for which in ["pix1", "pix2", "ipix1", "ipix2", "logratio",
                          "binw", "binv1", "binv2", "binlogratio"]:
                assert np.array_equal(g.array(s, which), o.array(s, which)), \
                    (s, which)


def test_data2_binning_bit_exact():
    import pyoracle2d as O2
    rng = np.random.default_rng(8)
    d1, d2 = rng.normal(size=30000), rng.normal(size=30000) * 3
    gb, gi = _lib.data2_bin(d1, d2, 77, 55)
    ob, oi = O2.data2_bin(d1, d2, 77, 55)
    assert np.array_equal(gb, ob) and np.array_equal(gi, oi)
    g1, g2, gm = _lib.data2_mean_subtract(d1, d2)
    o1, o2, om = O2.data2_mean_subtract(d1, d2)
    assert np.array_equal(g1, o1) and np.array_equal(g2, o2)
    assert np.array_equal(gm, om)


def test_acor_ar1_closed_form():
    """affineChainSet::acor (Goodman's ACOR, src/affineChainSet.cc:880-950)
    on AR(1) series x_t = rho x_{t-1} + e_t, whose integrated
    autocorrelation time is (1 + rho)/(1 - rho) and whose mean has standard
    error sqrt(tau var / n); short series are refused (L < minfac * maxlag
    = 50).  The reference's own fixture is the next test."""
    rng = np.random.default_rng(12)
    for rho in (0.0, 0.5, 0.9, 0.98):
        n = 1 << 20
        e = rng.normal(size=n)
        from scipy.signal import lfilter
        x = lfilter([1.0], [1.0, -rho], e)
        tau, mean, sigma = _lib.acor(x)
        want = (1 + rho) / (1 - rho)
        assert tau == pytest.approx(want, rel=0.08), (rho, tau, want)
        var = 1.0 / (1 - rho * rho)
        assert sigma == pytest.approx(np.sqrt(want * var / n), rel=0.08)
        assert abs(mean - x.mean()) < 1e-5 * (1 + abs(x.mean()))
    with pytest.raises(_lib.PofdError):
        _lib.acor(np.zeros(49))


def test_acor_reference_vector():
    """The reference's known answer (test/chain_test.cc:461-487): ACOR of
    the 8815-value series testdata/testvec.txt read as floats gives
    tau 130.052 +- 0.002, mean 5.82089 +- 0.0002, sigma 0.0134451 +- 0.0002
    ("values computed using original acor code")."""
    import os
    from conftest import REFDATA
    x = np.loadtxt(os.path.join(REFDATA, "testvec.txt"), dtype=np.float32)
    assert x.size == 8815
    tau, mean, sigma = _lib.acor(x)
    assert abs(tau - 130.052) < 0.002, tau
    assert abs(mean - 5.82089) < 0.0002, mean
    assert abs(sigma - 0.0134451) < 0.0002, sigma
