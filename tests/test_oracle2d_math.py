"""Mathematical pins of the 2D oracle where the reference's tests leave it
open: transforms against numpy, the log-normal solver against scipy,
Campbell's theorem for the P(D) moments, and the Gaussian limit."""
import math

import numpy as np
import pytest
from scipy import optimize

import pyoracle2d as O2
from pofd_affine_amd import synth2d
from oracle2d_side import build_oracle2d, make_inputs2d


@pytest.mark.parametrize("n", [8, 64, 256])
def test_rfft2_irfft2_match_numpy(n):
    x = np.random.default_rng(n).normal(size=(n, n))
    X = O2.rfft2(x)
    np.testing.assert_allclose(X, np.fft.rfft2(x), rtol=0,
                               atol=1e-12 * np.abs(X).max())
    y = O2.irfft2(X, n)
    np.testing.assert_allclose(y, x * n * n, rtol=0, atol=1e-10 * n * n)


def test_lognormal_solver_matches_brentq():
    """getMaxFlux band 2 = S1_max * exp(u), u the root above the mode where
    the log-normal density in u falls to exp(-nsig^2/2) of its peak."""
    kp, sp, op = synth2d.positions()
    m = O2.LogNormalModel(kp, sp, op)
    m.set_params(synth2d.true_theta())
    mx = m.maxflux()
    mu = m.offset(kp[-1])
    sg = m.sigma(kp[-1])
    sig2 = sg * sg
    mode = mu - sig2

    def dens(u):
        return math.exp(-0.5 * (u - mu) ** 2 / sig2) * math.exp(-u) / sg

    target = dens(mode) * math.exp(-0.5 * 9.0)
    u = optimize.brentq(lambda v: dens(v) - target, mode, mode + 20 * sg,
                        xtol=1e-14)
    assert mx[0] == kp[-1]
    assert mx[1] / kp[-1] == pytest.approx(math.exp(u), rel=2e-3)


@pytest.fixture(scope="module")
def small2d():
    inp = make_inputs2d(npix_side=120, seed=5)
    th = synth2d.true_theta()
    return inp, th


def test_pd2d_moments_campbell(small2d):
    """Compound Poisson + Gaussian noise per band: the normalised 2D P(D)
    has mean 0 after the un-shift and marginal variances var_noi + sigma^2
    (the R moments of getRStats)."""
    inp, th = small2d
    n = 512  # at 256 the negative side is under 3 bins: findSplitPoint fails
    L = build_oracle2d(inp, n, th, nedge=64)
    ip = L.initpd(0, th)
    pd, info = L.getpd(0, 0, 1.0, 1.0, setlog=False)
    x1 = info["minflux1"] + np.arange(n) * info["dflux1"]
    x2 = info["minflux2"] + np.arange(n) * info["dflux2"]
    w = pd * info["dflux1"] * info["dflux2"]
    p1, p2 = w.sum(1), w.sum(0)
    m1, m2 = (x1 * p1).sum() / p1.sum(), (x2 * p2).sum() / p2.sum()
    v1 = ((x1 - m1) ** 2 * p1).sum() / p1.sum()
    v2 = ((x2 - m2) ** 2 * p2).sum() / p2.sum()
    assert abs(m1) < 2 * info["dflux1"] and abs(m2) < 2 * info["dflux2"]
    assert v1 == pytest.approx(ip["var1"] + synth2d.SIGMA1 ** 2, rel=2e-2)
    assert v2 == pytest.approx(ip["var2"] + synth2d.SIGMA2 ** 2, rel=2e-2)


def test_pd2d_negligible_R_is_gaussian(small2d):
    """Counts 1e-30 below the fiducial ones: R vanishes, the P(D) is the
    product of the two instrument Gaussians, and logL is their Gaussian
    log likelihood (to the bilinear-interpolation error of log2 P)."""
    inp, _ = small2d
    th = synth2d.true_theta().copy()
    th[:len(synth2d.KNOTS_MJY)] -= 30.0
    rng = np.random.default_rng(3)
    s1, s2 = synth2d.SIGMA1, synth2d.SIGMA2
    d1 = rng.normal(0, s1, 20000)
    d2 = rng.normal(0, s2, 20000)
    d1 -= d1.mean()
    d2 -= d2.mean()
    inp2 = dict(inp, d1=d1, d2=d2)
    n = 512
    L = build_oracle2d(inp2, n, th, nedge=64)
    ll, rc = L.loglike(th)
    assert rc == 0
    info = L.beamset_info(0)
    gauss = np.sum(-0.5 * np.log(2 * np.pi * s1 * s1) - 0.5 * (d1 / s1) ** 2
                   - 0.5 * np.log(2 * np.pi * s2 * s2) - 0.5 * (d2 / s2) ** 2)
    curr = ll + info["like_offset0"]
    df1 = (info["maxR1"] - info["minR1"]) / (n - 1)
    df2 = (info["maxR2"] - info["minR2"]) / (n - 1)
    tol = d1.size * (df1 ** 2 / (8 * s1 * s1) + df2 ** 2 / (8 * s2 * s2)) \
        + 1e-6 * abs(gauss)
    assert abs(curr - gauss) < tol


def test_data2_binning_centres():
    rng = np.random.default_rng(1)
    d1, d2 = rng.normal(size=5000), rng.normal(size=5000) * 2
    bins, info = O2.data2_bin(d1, d2, 40, 30)
    assert bins.sum() == 5000
    assert info[0] == d1.min() and info[2] == d2.min()
    assert info[0] + 39 * info[1] == pytest.approx(d1.max(), rel=1e-14)
    i1 = ((d1 - info[0]) / info[1] + 0.5).astype(int)
    i2 = ((d2 - info[2]) / info[3] + 0.5).astype(int)
    ref = np.zeros((40, 30), dtype=np.uint32)
    np.add.at(ref, (i1, i2), 1)
    assert np.array_equal(ref, bins)


def test_r2d_grid_and_edge_integrals(small2d):
    """R on the grid (initR with setEdge, src/PDFactoryDouble.cc:376-551):
    away from the axes it is the array getR (pinned by the reference's golden
    values) times dflux1 dflux2; the first row, first column and corner are
    the reference's trapezoid averages of R over the cells [0, dflux] (log
    spaced in band 1, linear in band 2).  With many edge points these must
    converge to the exact cell averages, computed here by dense quadrature
    of getR -- which a wrong edge normalisation, edge grid or axis fails."""
    from scipy import integrate
    inp, th = small2d
    n = 512
    L = build_oracle2d(inp, n, th, nedge=4096)
    ip = L.initpd(0, th)
    R = ip["rvals"]
    d1f, d2f = ip["dflux1"], ip["dflux2"]
    assert int(ip["wrap1"]) == n - 1  # positive beams: no wrapped fluxes
    kp, sp, op = synth2d.positions()
    m = O2.LogNormalModel(kp, sp, op)
    m.set_params(th[:len(kp) + len(sp) + len(op)])
    bm = O2.DoubleBeam(inp["beam1"], inp["beam2"], inp["pixsize"],
                       inp["minbeamval"], inp["nbeamhist"])
    rng = np.random.default_rng(2)
    rows = rng.choice(np.arange(1, n), 12, replace=False)
    ref = m.getR(rows * d1f, np.arange(1, n) * d2f, bm) * d1f * d2f
    np.testing.assert_allclose(R[rows, 1:], ref, rtol=1e-12, atol=0)
    lo1, lo2 = 1e-9 * d1f, 1e-9 * d2f  # lowEdgeRMult
    cell = d1f * d2f
    # first row: band-1 average over [lo1, dflux1] of R(f1, f2_j), dense in
    # ln f1 (the integrand has narrow log-normal features)
    js = np.array(list(np.argsort(R[0, 1:])[::-1][:200:25] + 1))
    u = np.linspace(math.log(lo1), math.log(d1f), 200001)
    G = m.getR(np.exp(u), js * d2f, bm) * np.exp(u)[:, None]
    row0 = np.trapezoid(G, u, axis=0) / (d1f - lo1) * cell
    np.testing.assert_allclose(R[0, js], row0, rtol=2e-3)
    # first column: band-2 average over [lo2, dflux2] of R(f1_i, f2)
    iis = list(np.argsort(R[1:, 0])[::-1][:200:25] + 1)
    for i in iis:
        v, _ = integrate.quad(lambda y: m.getR1(i * d1f, y, bm), lo2, d2f,
                              limit=400, epsrel=1e-10)
        assert R[i, 0] == pytest.approx(v / (d2f - lo2) * cell, rel=2e-3), i
    # corner: both averages
    u = np.linspace(math.log(lo1), math.log(d1f), 40001)
    y = np.linspace(lo2, d2f, 401)
    G = m.getR(np.exp(u), y, bm) * np.exp(u)[:, None]
    c00 = np.trapezoid(np.trapezoid(G, y, axis=1), u) / (
        (d1f - lo1) * (d2f - lo2)) * cell
    # (slowest to converge: linear edge grid in band 2 down to f2 -> 0)
    assert R[0, 0] == pytest.approx(c00, rel=1e-2)


def test_split_points_2d_properties(small2d):
    """findSplitPoint on both marginals (src/PDFactoryDouble.cc:784-975) has
    no reference fixture: after the rotation the marginals start and end in
    their tails (below f2 = 1.5e-5 of the peak) and the P(D) has zero mean
    on both axes."""
    inp, th = small2d
    n = 512
    L = build_oracle2d(inp, n, th, nedge=64)
    L.initpd(0, th, want_arrays=False)
    pd, info = L.getpd(0, 0, 1.0, 1.0, setlog=False)
    p1, p2 = pd.sum(1), pd.sum(0)
    for p in (p1, p2):
        assert max(p[0], p[-1]) < 1.5e-5 * p.max()
    x1 = info["minflux1"] + np.arange(n) * info["dflux1"]
    x2 = info["minflux2"] + np.arange(n) * info["dflux2"]
    assert abs(np.sum(x1 * p1) / p1.sum()) < 1e-3 * info["dflux1"]
    assert abs(np.sum(x2 * p2) / p2.sum()) < 1e-3 * info["dflux2"]
    assert info["split1"] * info["dflux1"] >= 2.25 * info["sgpos1"]
    assert info["split2"] * info["dflux2"] >= 2.25 * info["sgpos2"]


def test_splineint2d_matches_adaptive_quadrature():
    """splineInt(alpha, beta) (src/numberCountsDoubleLogNormal.cc:
    1549-1673): band-2 moments of the log-normal colour model are
    int S1^(a+b) exp(b mu + b^2 sigma^2/2) n1 dS1; checked against scipy's
    adaptive integration of the same integrand built from the oracle's
    golden-pinned band-1 counts, sigma(S1) and offset(S1), and against a
    direct 2D integral of the counts for band 2."""
    from scipy import integrate
    kp, sp, op = synth2d.positions()
    m = O2.LogNormalModel(kp, sp, op)
    th = synth2d.true_theta()
    m.set_params(th[:len(kp) + len(sp) + len(op)])
    brk = np.log(np.concatenate([kp[1:-1], sp, op]))
    brk = brk[(brk > math.log(kp[0])) & (brk < math.log(kp[-1]))]
    for a, b in ((0, 0), (1, 0), (0, 1), (2, 0), (0, 2)):
        def f(u):
            s1 = math.exp(u)
            e = b * m.offset(s1) + 0.5 * b * b * m.sigma(s1) ** 2
            return s1 ** (a + b + 1) * math.exp(e) * m.band1_counts(s1)
        ref, _ = integrate.quad(f, math.log(kp[0]), math.log(kp[-1]),
                                epsrel=1e-13, limit=500, points=brk)
        assert m.splineint(a, b) == pytest.approx(ref, rel=1e-10), (a, b)
    # band-2 flux per area directly: int int S2 n(S1, S2) dS2 dS1
    # (the counts vanish outside [minknot, maxknot): keep the grid just
    # inside both ends)
    s1g = np.exp(np.linspace(math.log(kp[0]) + 1e-12,
                             math.log(kp[-1]) - 1e-12, 4001))
    inner = []
    for s1 in s1g:
        v, _ = integrate.quad(lambda w: math.exp(w) * math.exp(w) *
                              m.counts(s1, math.exp(w)),
                              math.log(s1) - 6, math.log(s1) + 6,
                              epsrel=1e-10, limit=200)
        inner.append(v * s1)
    direct = np.trapezoid(inner, np.log(s1g))
    assert m.splineint(0, 1) == pytest.approx(direct, rel=1e-5)
