"""The oracle's PD::getLogLike and PDDouble::getLogLike against closed forms.

With log2 P(D) linear in flux (1D) or bilinear in the two fluxes (2D), the
reference's linear (bilinear) interpolation of log2 P(D)
(src/PD.cc:557-677, src/PDDouble.cc:976-1222) is exact inside the grid, so
the log likelihood is ln 2 times the sum of that function at the data
points; outside, each edge rule clamps to a known grid value.  Points are
placed between the two candidate edges (maxflux - dflux and maxflux in 1D)
so that a wrong comparison (< for <=, maxflux for maxflux - dflux) changes
the answer.  The P(D) stored linearly (logflat false) must give the same
sums, since log2 of 2^f is f.
"""
import numpy as np
import pytest

import pyoracle as O
import pyoracle2d as O2

LN2 = np.log(2.0)


def _grid1d():
    n, minflux, dflux = 64, -0.01, 0.0005
    x = minflux + dflux * np.arange(n)
    a, b = -3.0, 40.0
    return n, minflux, dflux, x, (lambda v: a + b * v)


@pytest.mark.parametrize("islog", [True, False])
def test_unbinned_linear_log_pd(islog):
    n, minflux, dflux, x, f = _grid1d()
    maxflux = x[-1]
    lpd = f(x)
    pd = lpd if islog else 2.0 ** lpd
    rng = np.random.default_rng(3)
    inside = rng.uniform(minflux + 1e-9, maxflux - dflux - 1e-9, 500)
    # edges of the unbinned sum: x <= minflux -> pd[0], x >= maxflux - dflux
    # -> pd[n-1] (a point just above maxflux - dflux is clamped although
    # it is inside the grid)
    edge = np.array([minflux, minflux - 0.003, maxflux - dflux,
                     maxflux - 0.4 * dflux, maxflux + 0.01])
    want_edge = np.array([f(minflux), f(minflux), f(maxflux), f(maxflux),
                          f(maxflux)])
    data = np.concatenate([inside, edge])
    v, _ = O.pd_loglike_unbinned(pd, minflux, dflux, data, islog=islog)
    want = LN2 * (np.sum(f(inside)) + np.sum(want_edge))
    assert v == pytest.approx(want, rel=1e-13, abs=1e-11)


@pytest.mark.parametrize("islog", [True, False])
def test_binned_linear_log_pd(islog):
    n, minflux, dflux, x, f = _grid1d()
    maxflux = x[-1]
    lpd = f(x)
    pd = lpd if islog else 2.0 ** lpd
    # bin centres from below minflux to above maxflux; the binned sum clamps
    # only for c < minflux and c > maxflux - dflux, interpolating at the
    # points themselves
    nb = 101
    c0 = minflux - 0.0031
    dl = (maxflux + 0.0017 - c0) / (nb - 1)
    rng = np.random.default_rng(4)
    bins = rng.integers(0, 7, nb).astype(np.uint32)
    c = c0 + dl * np.arange(nb)
    val = np.where(c < minflux, f(minflux),
                   np.where(c > maxflux - dflux, f(maxflux), f(c)))
    want = LN2 * np.sum(bins * val)
    v = O.pd_loglike_binned(pd, minflux, dflux, bins, c0, dl, islog=islog)
    assert v == pytest.approx(want, rel=1e-13)
    assert np.any((c > maxflux - dflux) & (c <= maxflux) & (bins > 0))


def _grid2d(n1=40, n2=40):
    m1, d1, m2, d2 = -0.01, 0.001, -0.02, 0.0015
    x1 = m1 + d1 * np.arange(n1)
    x2 = m2 + d2 * np.arange(n2)

    def f(u, v):
        return -2.0 + 30.0 * u - 12.0 * v + 400.0 * u * v
    return m1, d1, m2, d2, x1, x2, f


@pytest.mark.parametrize("islog", [True, False])
def test_2d_unbinned_bilinear_log_pd(islog):
    m1, d1, m2, d2, x1, x2, f = _grid2d()
    M1, M2 = x1[-1], x2[-1]
    lpd = f(x1[:, None], x2[None, :])
    pd = lpd if islog else 2.0 ** lpd
    rng = np.random.default_rng(5)
    u = rng.uniform(m1 + 1e-9, M1 - 1e-9, 400)
    v = rng.uniform(m2 + 1e-9, M2 - 1e-9, 400)
    # off-grid points clamp each coordinate to its edge (the 2D edges are
    # minflux and maxflux, not maxflux - dflux)
    eu = np.array([m1 - 0.005, M1 + 0.002, 0.3 * M1, m1 - 0.1, M1 + 0.1])
    ev = np.array([0.2 * M2, m2 - 0.003, M2 + 0.01, m2 - 0.1, M2 + 0.4])
    du = np.concatenate([u, eu])
    dv = np.concatenate([v, ev])
    cu = np.clip(du, m1, M1)
    cv = np.clip(dv, m2, M2)
    # an off-grid coordinate takes its grid edge; the other one is then not
    # interpolated but truncated to its cell's lower node
    idx1 = np.floor((cu - m1) / d1)
    idx2 = np.floor((cv - m2) / d2)
    off1 = (du <= m1) | (du >= M1)
    off2 = (dv <= m2) | (dv >= M2)
    gu = np.where(off2 & ~off1, m1 + d1 * idx1, cu)
    gv = np.where(off1 & ~off2, m2 + d2 * idx2, cv)
    want = LN2 * np.sum(f(gu, gv))
    got = O2.lib().orc_pd2_loglike_unbinned(
        40, 40, O2._dp(O2._f64(pd)), m1, d1, m2, d2, int(islog),
        O2._dp(O2._f64(du)), O2._dp(O2._f64(dv)), du.size)
    assert got == pytest.approx(want, rel=1e-12)


@pytest.mark.parametrize("islog", [True, False])
def test_2d_binned_bilinear_log_pd(islog):
    m1, d1, m2, d2, x1, x2, f = _grid2d()
    M1, M2 = x1[-1], x2[-1]
    lpd = f(x1[:, None], x2[None, :])
    pd = lpd if islog else 2.0 ** lpd
    nb1, nb2 = 23, 31
    c01, dl1 = m1 - 0.004, (M1 + 0.003 - (m1 - 0.004)) / (nb1 - 1)
    c02, dl2 = m2 - 0.006, (M2 + 0.002 - (m2 - 0.006)) / (nb2 - 1)
    rng = np.random.default_rng(6)
    bins = rng.integers(0, 5, (nb1, nb2)).astype(np.uint32)
    c1 = c01 + dl1 * np.arange(nb1)
    c2 = c02 + dl2 * np.arange(nb2)
    C1, C2 = np.meshgrid(c1, c2, indexing="ij")
    off1 = (C1 < m1) | (C1 > M1)
    off2 = (C2 < m2) | (C2 > M2)
    cu, cv = np.clip(C1, m1, M1), np.clip(C2, m2, M2)
    gu = np.where(off2 & ~off1, m1 + d1 * np.floor((cu - m1) / d1), cu)
    gv = np.where(off1 & ~off2, m2 + d2 * np.floor((cv - m2) / d2), cv)
    want = LN2 * np.sum(bins * f(gu, gv))
    b = np.ascontiguousarray(bins)
    got = O2.lib().orc_pd2_loglike_binned(
        40, 40, O2._dp(O2._f64(pd)), m1, d1, m2, d2, int(islog),
        b.ctypes.data_as(O2.U32P), nb1, nb2, c01, dl1, c02, dl2)
    assert got == pytest.approx(want, rel=1e-12)
