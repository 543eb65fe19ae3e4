"""The CPU oracle against library routines, closed forms and brute force.

These pin the parts of the oracle that the reference's own tests leave open:
the restated GSL natural spline (scipy), the restated FFTW transforms (numpy's
pocketfft), the data binning (the bin-centre property), and the P(D)
pipeline (Campbell's theorem cumulants, a Gaussian closed form when R is
negligible, and an O(n^2) direct DFT at small n).
"""
import math

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

import pyoracle as O
from pofd_affine_amd import synth


def test_cspline_matches_scipy_natural():
    rng = np.random.default_rng(3)
    x = np.cumsum(rng.uniform(0.1, 1.0, 40))
    y = rng.normal(size=40)
    c = O.cspline_init(x, y)
    ref = CubicSpline(x, y, bc_type="natural")
    # GSL's c is half the second derivative
    np.testing.assert_allclose(c, 0.5 * ref(x, 2), rtol=1e-10, atol=1e-12)
    xv = np.linspace(x[0], x[-1], 777)
    np.testing.assert_allclose(O.cspline_eval(x, y, c, xv), ref(xv),
                               rtol=1e-12, atol=1e-12)


def test_cspline_linear_data_is_exact():
    x = np.log2(np.array([0.002, 0.005, 0.01, 0.02, 0.045]))
    y = 3.0 - 2.2 * x
    c = O.cspline_init(x, y)
    assert np.max(np.abs(c)) < 1e-12
    xv = np.linspace(x[0], x[-1], 101)
    np.testing.assert_allclose(O.cspline_eval(x, y, c, xv), 3.0 - 2.2 * xv,
                               rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("n", [8, 64, 1024, 131072])
def test_rfft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n)
    got = O.rfft(x)
    ref = np.fft.rfft(x)
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref)) * math.log2(n)


@pytest.mark.parametrize("n", [8, 64, 1024, 131072])
def test_irfft_matches_numpy_unnormalised(n):
    rng = np.random.default_rng(n + 1)
    X = rng.normal(size=n // 2 + 1) + 1j * rng.normal(size=n // 2 + 1)
    X[0] = X[0].real
    X[-1] = X[-1].real
    got = O.irfft(X, n)
    ref = np.fft.irfft(X, n) * n
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref)) * math.log2(n)
    # the imaginary parts of the DC and Nyquist terms are ignored
    Y = X.copy()
    Y[0] += 5j
    Y[-1] -= 3j
    np.testing.assert_allclose(O.irfft(Y, n), got, rtol=0, atol=1e-12 * n)


def test_data_binning_bin_centres():
    rng = np.random.default_rng(7)
    d = rng.normal(0.0, 0.01, 100001)
    bins, c0, dl = O.data_bin(d, 1000)
    assert bins.sum() == d.size
    assert c0 == d.min()
    assert dl == pytest.approx((d.max() - d.min()) / 999, rel=1e-15)
    # every datum lies within half a bin of its bin centre: recount with an
    # independent nearest-centre assignment (differs only at exact ties)
    idx = np.rint((d - c0) / dl).astype(np.int64)
    ref = np.bincount(idx, minlength=1000)
    assert np.abs(ref - bins.astype(np.int64)).sum() <= 2
    assert bins[0] >= 1 and bins[-1] >= 1


def test_getR_power_law_closed_form():
    """Collinear log knots give a pure power law n = K S^-g, so for a
    histogrammed beam R(x) = pix^2 sum_j w_j eta_j (x eta_j)^-g exactly."""
    # knots at powers of two and float32-exact values: collinear in
    # (log2 S, log2 n) up to one rounding of the log2(10) scaling
    kp = 2.0 ** np.array([-10.0, -8.0, -6.0, -4.0, -2.0])
    kv = np.array([7.0, 6.25, 5.5, 4.75, 4.0], dtype=np.float32)
    m = O.Knots(kp, kv)
    b2, ps = synth.gaussian_beam()
    b = O.Beam(b2.ravel(), ps, 1e-6, 120)
    w = b.array("posweights")
    ieta = b.array("poshistval")
    x = np.geomspace(2e-4, 0.3, 60)
    got = m.getR(x, b)
    lk = np.log2(10.0) * kv.astype(np.float64)
    gg = -(lk[-1] - lk[0]) / (np.log2(kp[-1]) - np.log2(kp[0]))
    K = 2.0 ** (lk[0] + gg * np.log2(kp[0]))
    ref = np.zeros_like(x)
    for i, xi in enumerate(x):
        s = xi * ieta
        ok = (s >= kp[0]) & (s < kp[-1])
        ref[i] = np.sum(w[ok] * ieta[ok] * K * s[ok] ** (-gg))
    ref *= (ps / 3600.0) ** 2
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)


def _small_setup(n, ninterp, sigma, seed=5, nside=120, vals=None,
                 exp_conf=0.0, data=None):
    b2, ps = synth.gaussian_beam()
    beam = O.Beam(b2.ravel(), ps, 1e-6, 120)
    d = synth.make_map(nside, sigma, seed=seed)[0] if data is None else data
    L = O.Like(n, ninterp, synth.knot_positions())
    b = L.add_beamset(beam)
    L.add_field(b, pix=d, sigma_base=sigma)
    th = synth.true_theta() if vals is None else vals
    L.finalize(th, exp_conf=exp_conf)
    return L, th, d


def test_pd_cumulants_campbell():
    """Compound Poisson + Gaussian noise: the normalised P(D) has variance
    var_noi + sigma^2 and third cumulant sum x^3 R dx (Campbell).  A large
    expected confusion noise widens the R range (calcLike.cc:361-364) so the
    P(D) tail does not alias around the transform."""
    sigma = 0.0025
    L, th, _ = _small_setup(32768, 512, sigma, exp_conf=0.2)
    ip = L.initpd(0, th)
    pd = L.getpd(0, 0, 1.0, setlog=False)
    p = pd["pd"]
    x = pd["minflux"] + np.arange(p.size) * pd["dflux"]
    w = p * pd["dflux"]
    m1 = np.sum(x * w)
    k2 = np.sum((x - m1) ** 2 * w)
    k3 = np.sum((x - m1) ** 3 * w)
    assert abs(m1) < 1e-6
    assert k2 == pytest.approx(ip["var_noi"] + sigma * sigma, rel=1e-8)
    dflux = ip["dflux"]
    xr = np.arange(ip["rvals"].size) * dflux
    k3_campbell = np.sum(xr ** 3 * ip["rvals"])  # rvals already carry dflux
    assert k3 == pytest.approx(k3_campbell, rel=1e-6)


def test_pd_negligible_R_is_gaussian():
    """With counts 1e-30 below the fiducial ones R vanishes, the P(D) is the
    instrument Gaussian and logL is the Gaussian log likelihood."""
    sigma = 0.0025
    vals = synth.true_theta().copy()
    vals[:8] -= 30.0
    d = np.random.default_rng(11).normal(0.0, sigma, 20000)
    d -= d.mean()
    L, th, d = _small_setup(32768, 256, sigma, vals=vals, data=d)
    ll = L.loglike(th)
    like_offset = -0.5 * d.size * (math.log(2 * math.pi * sigma * sigma) + 1)
    gauss = np.sum(-0.5 * math.log(2 * math.pi * sigma * sigma)
                   - 0.5 * (d / sigma) ** 2)
    # lerp of log2 P between grid points: error <= dflux^2/(8 sigma^2) per
    # pixel in natural-log units
    info = L.beamset_info(0)
    dflux = info["maxRFlux"] / (32768 - 1)
    tol = d.size * dflux * dflux / (8 * sigma * sigma) + 1e-6 * abs(gauss)
    assert abs((ll + like_offset) - gauss) < tol


def test_pd_pipeline_direct_dft_small_n():
    """At n = 1024 the FFT-based initPD/getPD matches the same pointwise
    operations done with an O(n^2) extended-precision direct DFT."""
    n, sigma = 1024, 0.004
    L, th, _ = _small_setup(n, 128, sigma, nside=60)
    ip = L.initpd(0, th)
    r = ip["rvals"].astype(np.longdouble)
    k = np.arange(n // 2 + 1)
    j = np.arange(n)
    ang = (-2 * np.pi * np.outer(k, j) / n).astype(np.longdouble)
    Rh = (r * np.cos(ang)).sum(1) + 1j * (r * np.sin(ang)).sum(1).astype(
        np.float64)
    Rh = np.array(Rh, dtype=np.complex128)
    assert ip["r0"] == pytest.approx(Rh[0].real, rel=1e-13)
    iflux = 2 * np.pi / (n * ip["dflux"])
    ph = Rh.imag - ip["shift"] * iflux * k
    pr = np.exp(Rh.real - Rh[0].real) * (np.cos(ph) + 1j * np.sin(ph))
    pr[0] = 1.0
    np.testing.assert_allclose(ip["prtrans"], pr, rtol=0, atol=1e-12)
    pv = pr * np.exp(-0.5 * sigma * sigma * (iflux * k) ** 2)
    pv[0] = 1.0
    full = np.concatenate([pv, np.conj(pv[1:-1][::-1])])
    full[n // 2] = pv[-1].real
    pofd = np.real(np.fft.ifft(full)) * n
    got = L.getpd(0, 0, 1.0, setlog=False)["pofd_raw"]
    ref = np.where(pofd < 0, 0.0, pofd)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * ref.max())


# ---- negative beams: the wrapped R grid (src/PDFactory.cc:361-414, 541-620)
def _negbeam_setup(n, ninterp, sigma, exp_conf=0.0, nhist=120, seed=7):
    import os
    from conftest import REFDATA
    from pofd_affine_amd import fitsio
    pix, ps = fitsio.read_beam(os.path.join(REFDATA,
                                            "band12_beam_f100.fits[0]"))
    beam = O.Beam(pix, ps, 1e-6, nhist)
    d = np.random.default_rng(seed).normal(0.0, sigma, 20000)
    d -= d.mean()
    L = O.Like(n, ninterp, synth.knot_positions())
    b = L.add_beamset(beam)
    L.add_field(b, pix=d, sigma_base=sigma)
    th = synth.true_theta()
    L.finalize(th, exp_conf=exp_conf)
    return L, th, beam


def _grid_flux(ip, n):
    i = np.arange(n)
    wr, dflux = ip["wrapRidx"], ip["dflux"]
    return np.where(i <= wr, i * dflux, (i - wr - 1) * dflux + ip["minflux_R"])


def test_negative_beam_wrapped_grid_is_getR():
    """Grid points above wrapRidx hold R at the negative fluxes
    (i - wrapRidx - 1) dflux + minflux_R; on both sides the filled values are
    the spline-interpolated array getR (itself pinned by the reference's
    negative-beam golden vectors) at those fluxes."""
    n = 32768
    L, th, beam = _negbeam_setup(n, 2048, 0.0025)
    ip = L.initpd(0, th)
    wr = ip["wrapRidx"]
    assert ip["minflux_R"] < 0 and 0 < wr < n - 1
    # the wrapped fluxes continue the positive ones around the circle
    x = _grid_flux(ip, n)
    assert x[wr + 1] == pytest.approx(x[wr] + ip["dflux"] - n * ip["dflux"],
                                      rel=1e-9)
    m = O.Knots(synth.knot_positions(), th[:len(synth.knot_positions())])
    r = ip["rvals"] / ip["dflux"]
    ref = m.getR(x, beam)
    for side in (x > 0, x < 0):
        sel = side & (ref > 1e-8 * ref[side].max())
        assert sel.sum() > 100
        rel = np.abs(r[sel] - ref[sel]) / ref[sel]
        # the tails where beam entries leave the knot range interpolate worst
        assert np.median(rel) < 1e-4 and np.percentile(rel, 90) < 2e-2, (
            np.median(rel), np.percentile(rel, 90))
    # a one-cell shift of the wrapped fluxes would be caught
    neg = (x < 0) & (ref > 1e-8 * ref[x < 0].max())
    shifted = m.getR(x[neg] + ip["dflux"], beam)
    ok = shifted > 0
    assert np.median(np.abs(r[neg][ok] - shifted[ok]) / shifted[ok]) > 1e-3


def test_negative_beam_pd_cumulants_campbell():
    """With a beam of both signs the normalised P(D) has variance
    sum x^2 R dx + sigma^2 and third cumulant sum x^3 R dx over the wrapped
    grid fluxes (Campbell)."""
    sigma = 0.0025
    n = 32768
    L, th, _ = _negbeam_setup(n, 512, sigma, exp_conf=0.2)
    ip = L.initpd(0, th)
    pd = L.getpd(0, 0, 1.0, setlog=False)
    p = pd["pd"]
    x = pd["minflux"] + np.arange(p.size) * pd["dflux"]
    w = p * pd["dflux"]
    m1 = np.sum(x * w)
    k2 = np.sum((x - m1) ** 2 * w)
    k3 = np.sum((x - m1) ** 3 * w)
    assert abs(m1) < 1e-6
    xr = _grid_flux(ip, n)
    k2_campbell = np.sum(xr ** 2 * ip["rvals"]) + sigma * sigma
    k3_campbell = np.sum(xr ** 3 * ip["rvals"])
    assert k2 == pytest.approx(k2_campbell, rel=1e-8)
    assert k3 == pytest.approx(k3_campbell, rel=1e-6)
    # the negative lobe matters: dropping it changes the third cumulant
    pos = xr > 0
    assert abs(np.sum(xr[pos] ** 3 * ip["rvals"][pos]) - k3_campbell) > \
        1e-4 * abs(k3_campbell)


def test_split_point_properties():
    """findSplitPoint (src/PDFactory.cc:691-921) has no reference fixture.
    Its contract: the unrotated P(D) is cut where it is below f2 = 1e-5 of
    its peak, at least n1 = 2.5 sigma above zero flux, so the rotated,
    normalised P(D) starts and ends in its tails and has zero mean."""
    sigma = 0.0025
    L, th, _ = _small_setup(32768, 512, sigma)
    for w in range(3):
        t = th.copy()
        t[:8] += 0.05 * w
        L.initpd(0, t)
        pd = L.getpd(0, 0, 1.0, setlog=False)
        raw, p = pd["pofd_raw"], pd["pd"]
        s = int(pd["split"])
        assert raw[s - 1] < 1e-5 * raw.max() or raw[s] < 1e-5 * raw.max()
        assert s * pd["dflux"] >= 2.5 * sigma
        assert max(p[0], p[-1]) < 1e-5 * p.max()
        x = pd["minflux"] + np.arange(p.size) * pd["dflux"]
        assert abs(np.sum(x * p) / np.sum(p)) < 1e-3 * pd["dflux"]


def test_splineint_matches_adaptive_quadrature():
    """getNS / getFluxPerArea / getFluxSqPerArea (splineInt,
    src/numberCountsKnotsSpline.cc:300-343): the oracle's Gauss-Legendre
    sums equal scipy's adaptive QUADPACK integration of S^alpha n(S), with
    n(S) the golden-pinned number counts, to 1e-10 (the reference's own
    QAG-41 stops at relative 1e-5)."""
    from scipy import integrate
    import ctypes
    kp = synth.knot_positions()
    rng = np.random.default_rng(8)
    for t in range(4):
        vals = synth.true_theta()[:len(kp)] + rng.normal(0, 0.3, len(kp))
        m = O.Knots(kp, vals.astype(np.float32))
        nc = lambda s: O.lib().orc_knots_number_counts(ctypes.byref(m.s), s)
        for a in (0, 1, 2):
            ref, _ = integrate.quad(lambda u: math.exp((a + 1) * u) *
                                    nc(math.exp(u)), math.log(kp[0]),
                                    math.log(kp[-1]), epsrel=1e-13,
                                    limit=500, points=np.log(kp[1:-1]))
            assert m.splineint(a) == pytest.approx(ref, rel=1e-10), (t, a)
