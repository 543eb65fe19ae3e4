"""ctypes front end of the 2D CPU oracle (oracle/pofd_oracle2d.c).

TEST INFRASTRUCTURE ONLY, like pyoracle.py: imported by tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py as the checker
or the reported CPU baseline, never as the product path.
"""
import ctypes as C

import numpy as np

import pyoracle as O

D = C.c_double
DP = C.POINTER(C.c_double)
U32P = C.POINTER(C.c_uint32)
I32P = C.POINTER(C.c_int32)
FP = C.POINTER(C.c_float)
VP = C.c_void_p
EP = C.POINTER(O.OrcErr)

_done = False


def lib():
    global _done
    L = O.lib()
    if _done:
        return L
    sig = {
        "orc_rfft2": (C.c_int, [C.c_uint, DP, DP]),
        "orc2_find_split_array": (C.c_int, [DP, C.c_uint, D, D, D,
                                            C.POINTER(C.c_uint), EP]),
        "orc_irfft2": (C.c_int, [C.c_uint, DP, DP]),
        "orc_dbeam_new": (VP, []),
        "orc_dbeam_free": (None, [VP]),
        "orc_dbeam_set": (C.c_int, [VP, DP, DP, C.c_size_t, D, D, EP]),
        "orc_dbeam_make_histogram": (C.c_int, [VP, C.c_uint32, EP]),
        "orc_dbeam_info": (None, [VP, DP]),
        "orc_dbeam_array": (DP, [VP, C.c_int, C.c_int, U32P]),
        "orc_ln_init": (None, [VP]),
        "orc_ln_set_positions": (C.c_int, [VP, C.c_uint, DP, C.c_uint, DP,
                                           C.c_uint, DP, EP]),
        "orc_ln_set_min_sigma": (C.c_int, [VP, D, EP]),
        "orc_ln_set_params": (C.c_int, [VP, C.c_uint, FP, EP]),
        "orc_ln_is_valid": (C.c_int, [VP]),
        "orc_ln_sigma": (D, [VP, D]),
        "orc_ln_offset": (D, [VP, D]),
        "orc_ln_counts": (D, [VP, D, D]),
        "orc_ln_band1_counts": (D, [VP, D]),
        "orc_ln_splineint": (D, [VP, D, D]),
        "orc_ln_maxflux": (C.c_int, [VP, DP, EP]),
        "orc_ln_rrange": (C.c_int, [VP, VP, DP, EP]),
        "orc_ln_getR1": (D, [VP, D, D, VP]),
        "orc_ln_getR": (None, [VP, C.c_uint, DP, C.c_uint, DP, VP, DP]),
        "orc_data2_minmax": (None, [DP, DP, C.c_size_t, DP]),
        "orc_data2_mean_subtract": (None, [DP, DP, C.c_size_t, DP]),
        "orc_data2_bin": (C.c_int, [DP, DP, C.c_size_t, C.c_uint32,
                                    C.c_uint32, U32P, DP]),
        "orc_pd2_integral": (D, [C.c_uint, C.c_uint, DP, D, D, C.c_int]),
        "orc_pd2_means": (None, [C.c_uint, C.c_uint, DP, D, D, D, D, C.c_int,
                                 C.c_int, DP]),
        "orc_pd2_loglike_unbinned": (D, [C.c_uint, C.c_uint, DP, D, D, D, D,
                                         C.c_int, DP, DP, C.c_size_t]),
        "orc_pd2_loglike_binned": (D, [C.c_uint, C.c_uint, DP, D, D, D, D,
                                       C.c_int, U32P, C.c_uint32, C.c_uint32,
                                       D, D, D, D]),
        "orc_like2_new": (VP, [C.c_uint, C.c_uint, C.c_int]),
        "orc_like2_free": (None, [VP]),
        "orc_like2_error": (EP, [VP]),
        "orc_like2_set_positions": (C.c_int, [VP, C.c_uint, DP, C.c_uint, DP,
                                              C.c_uint, DP, D]),
        "orc_like2_add_beamset": (C.c_int, [VP, VP]),
        "orc_like2_add_field": (C.c_int, [VP, C.c_int, DP, DP, C.c_size_t,
                                          U32P, C.c_uint32, C.c_uint32, DP,
                                          D, D, D]),
        "orc_like2_finalize_setup": (C.c_int, [VP, FP, C.c_uint, D, D]),
        "orc_like2_set_rrange": (C.c_int, [VP, C.c_int, DP]),
        "orc_like2_beamset_info": (C.c_int, [VP, C.c_int, DP]),
        "orc_like2_loglike": (C.c_int, [VP, FP, C.c_uint, DP]),
        "orc_like2_loglike_batch": (C.c_int, [VP, FP, C.c_uint, C.c_uint,
                                              C.c_uint, DP, I32P, C.c_uint]),
        "orc_like2_initpd": (C.c_int, [VP, C.c_int, FP, C.c_uint, DP, DP,
                                       DP]),
        "orc_like2_getpd": (C.c_int, [VP, C.c_int, C.c_int, D, D, C.c_int,
                                      DP, DP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _done = True
    return L


def _dp(a):
    return a.ctypes.data_as(DP)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _check(rc, err):
    if rc:
        raise O.OracleError(rc, err.text())


def find_split2(data, dflux, sigmapos, sigmaneg):
    """PDFactoryDouble::findSplitPoint (src/PDFactoryDouble.cc:784-975) on a
    marginal; raises OracleError where it throws."""
    a = _f64(data)
    out = C.c_uint(0)
    err = O.OrcErr()
    _check(lib().orc2_find_split_array(_dp(a), a.size, dflux, sigmapos,
                                       sigmaneg, C.byref(out), C.byref(err)),
           err)
    return int(out.value)


def rfft2(x):
    x = _f64(x)
    n = x.shape[0]
    out = np.zeros(2 * n * (n // 2 + 1))
    lib().orc_rfft2(n, _dp(x), _dp(out))
    out = out[0::2] + 1j * out[1::2]
    return out.reshape(n, n // 2 + 1)


def irfft2(X, n):
    X = np.ascontiguousarray(X, dtype=np.complex128)
    inter = np.empty(2 * X.size)
    inter[0::2] = X.real.ravel()
    inter[1::2] = X.imag.ravel()
    out = np.zeros(n * n)
    lib().orc_irfft2(n, _dp(inter), _dp(out))
    return out.reshape(n, n)


class DoubleBeam:
    """doublebeam::setBeams (+ makeHistogram), src/doublebeam.cc."""

    KEYS = ["hassign", "npix", "ishist", "nhist", "tot1", "tot2", "totsq1",
            "totsq2", "min1", "max1", "min2", "max2"]

    def __init__(self, b1, b2, pixsize, minval=1e-6, nhist=None):
        b1, b2 = _f64(np.ravel(b1)), _f64(np.ravel(b2))
        assert b1.size == b2.size
        self.ptr = lib().orc_dbeam_new()
        err = O.OrcErr()
        _check(lib().orc_dbeam_set(self.ptr, _dp(b1), _dp(b2), b1.size,
                                   float(pixsize), float(minval),
                                   C.byref(err)), err)
        if nhist:
            _check(lib().orc_dbeam_make_histogram(self.ptr, int(nhist),
                                                  C.byref(err)), err)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_dbeam_free(self.ptr)
            self.ptr = None

    def info(self):
        out = np.zeros(64)
        lib().orc_dbeam_info(self.ptr, _dp(out))
        sgn = [dict(zip(self.KEYS, out[12 * s:12 * s + 12])) for s in range(4)]
        return sgn, dict(pixsize=out[48], minval=out[49], nbins=out[50])

    def array(self, sgn, which):
        names = ["pix1", "pix2", "ipix1", "ipix2", "logratio", "binw",
                 "binv1", "binv2", "binlogratio"]
        n = C.c_uint32()
        p = lib().orc_dbeam_array(self.ptr, sgn, names.index(which),
                                  C.byref(n))
        if not p or n.value == 0:
            return np.zeros(0)
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy()


class OrcLN(C.Structure):
    K = 64
    _fields_ = [("nknots", C.c_uint), ("nsigma", C.c_uint),
                ("noffset", C.c_uint),
                ("knots", D * K), ("logknots", D * K), ("logknotvals", D * K),
                ("kc", D * K),
                ("sigmaknots", D * K), ("sigmavals", D * K), ("sc", D * K),
                ("offsetknots", D * K), ("offsetvals", D * K), ("oc", D * K),
                ("min_sigma", D), ("pos_loaded", C.c_int),
                ("vals_loaded", C.c_int), ("knots_valid", C.c_int),
                ("sigmas_valid", C.c_int), ("offsets_valid", C.c_int)]


class LogNormalModel:
    """numberCountsDoubleLogNormal restated."""

    def __init__(self, knots, sigmapos, offsetpos, min_sigma=0.0):
        self.m = OrcLN()
        lib().orc_ln_init(C.byref(self.m))
        kp, sp, op = _f64(knots), _f64(sigmapos), _f64(offsetpos)
        err = O.OrcErr()
        _check(lib().orc_ln_set_positions(C.byref(self.m), kp.size, _dp(kp),
                                          sp.size, _dp(sp), op.size, _dp(op),
                                          C.byref(err)), err)
        _check(lib().orc_ln_set_min_sigma(C.byref(self.m), float(min_sigma),
                                          C.byref(err)), err)
        self.nparams = kp.size + sp.size + op.size

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        err = O.OrcErr()
        _check(lib().orc_ln_set_params(C.byref(self.m), p.size,
                                       p.ctypes.data_as(FP), C.byref(err)),
               err)

    def is_valid(self):
        return bool(lib().orc_ln_is_valid(C.byref(self.m)))

    def sigma(self, f1):
        return lib().orc_ln_sigma(C.byref(self.m), f1)

    def offset(self, f1):
        return lib().orc_ln_offset(C.byref(self.m), f1)

    def counts(self, f1, f2):
        return lib().orc_ln_counts(C.byref(self.m), f1, f2)

    def band1_counts(self, f1):
        return lib().orc_ln_band1_counts(C.byref(self.m), f1)

    def splineint(self, alpha, beta):
        """int S1^(a+b) exp(b mu + b^2 sigma^2 / 2) n1(S1) dS1 (getNS,
        getFluxPerArea, getFluxSqPerArea of either band)."""
        return lib().orc_ln_splineint(C.byref(self.m), float(alpha),
                                      float(beta))

    def maxflux(self):
        out = np.zeros(2)
        err = O.OrcErr()
        _check(lib().orc_ln_maxflux(C.byref(self.m), _dp(out), C.byref(err)),
               err)
        return out

    def rrange(self, bm):
        out = np.zeros(4)
        err = O.OrcErr()
        _check(lib().orc_ln_rrange(C.byref(self.m), bm.ptr, _dp(out),
                                   C.byref(err)), err)
        return out

    def getR1(self, f1, f2, bm):
        return lib().orc_ln_getR1(C.byref(self.m), f1, f2, bm.ptr)

    def getR(self, f1, f2, bm):
        f1, f2 = _f64(f1), _f64(f2)
        R = np.zeros(f1.size * f2.size)
        lib().orc_ln_getR(C.byref(self.m), f1.size, _dp(f1), f2.size, _dp(f2),
                          bm.ptr, _dp(R))
        return R.reshape(f1.size, f2.size)


def data2_mean_subtract(d1, d2):
    d1, d2 = _f64(d1).copy(), _f64(d2).copy()
    m = np.zeros(2)
    lib().orc_data2_mean_subtract(_dp(d1), _dp(d2), d1.size, _dp(m))
    return d1, d2, m


def data2_bin(d1, d2, nb1, nb2):
    d1, d2 = _f64(d1), _f64(d2)
    bins = np.zeros(nb1 * nb2, dtype=np.uint32)
    info = np.zeros(4)
    rc = lib().orc_data2_bin(_dp(d1), _dp(d2), d1.size, nb1, nb2,
                             bins.ctypes.data_as(U32P), _dp(info))
    if rc:
        raise O.OracleError(rc, "fitsDataDouble::applyBinning failed")
    return bins.reshape(nb1, nb2), info


def pd2_integral(pd, dflux1, dflux2, islog=False):
    pd = _f64(pd)
    return lib().orc_pd2_integral(pd.shape[0], pd.shape[1], _dp(pd), dflux1,
                                  dflux2, int(islog))


def pd2_means(pd, minflux1, dflux1, minflux2, dflux2, islog=False,
              donorm=True):
    pd = _f64(pd)
    out = np.zeros(2)
    lib().orc_pd2_means(pd.shape[0], pd.shape[1], _dp(pd), minflux1, dflux1,
                        minflux2, dflux2, int(islog), int(donorm), _dp(out))
    return out


class Like2:
    """calcLikeDouble restated: beam sets of pixel-pair datasets."""

    def __init__(self, fftsize, nedge=256, edge_integ=True):
        self.ptr = lib().orc_like2_new(fftsize, nedge, int(edge_integ))
        if not self.ptr:
            raise ValueError("bad fftsize/nedge")
        self.n = fftsize
        self._keep = []

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_like2_free(self.ptr)
            self.ptr = None

    def _err(self, rc, what):
        if rc:
            e = lib().orc_like2_error(self.ptr).contents
            raise O.OracleError(rc, "%s: %s" % (what, e.text()))

    def set_positions(self, knots, sigmapos, offsetpos, min_sigma=0.0):
        kp, sp, op = _f64(knots), _f64(sigmapos), _f64(offsetpos)
        self.nmodel = kp.size + sp.size + op.size
        self._err(lib().orc_like2_set_positions(self.ptr, kp.size, _dp(kp),
                                                sp.size, _dp(sp), op.size,
                                                _dp(op), float(min_sigma)),
                  "set_positions")

    def add_beamset(self, dbeam):
        self._keep.append(dbeam)
        return lib().orc_like2_add_beamset(self.ptr, dbeam.ptr)

    def add_field(self, b, d1, d2, sigma_base1, sigma_base2, like_norm=1.0,
                  bins=None, bininfo=None):
        d1, d2 = _f64(d1), _f64(d2)
        if bins is not None:
            bins = np.ascontiguousarray(bins, dtype=np.uint32)
            nb1, nb2 = bins.shape
            info = _f64(bininfo)
            bp, ip = bins.ctypes.data_as(U32P), _dp(info)
        else:
            nb1 = nb2 = 0
            bp, ip = C.cast(None, U32P), C.cast(None, DP)
        rc = lib().orc_like2_add_field(self.ptr, b, _dp(d1), _dp(d2), d1.size,
                                       bp, nb1, nb2, ip, sigma_base1,
                                       sigma_base2, like_norm)
        if rc < 0:
            raise O.OracleError(rc, "add_field failed")
        return rc

    def finalize(self, theta, exp_conf1=0.0, exp_conf2=0.0):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        self._err(lib().orc_like2_finalize_setup(self.ptr, t.ctypes.data_as(FP),
                                                 t.size, exp_conf1, exp_conf2),
                  "finalize")

    def set_rrange(self, b, rng):
        r = _f64(rng)
        self._err(lib().orc_like2_set_rrange(self.ptr, b, _dp(r)), "rrange")

    def beamset_info(self, b):
        out = np.zeros(12)
        self._err(lib().orc_like2_beamset_info(self.ptr, b, _dp(out)), "info")
        keys = ["minR1", "maxR1", "minR2", "maxR2", "maxsig1", "maxsig2",
                "mind1", "maxd1", "mind2", "maxd2", "nfields", "like_offset0"]
        return dict(zip(keys, out))

    def loglike(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        v = D()
        rc = lib().orc_like2_loglike(self.ptr, t.ctypes.data_as(FP), t.size,
                                     C.byref(v))
        return v.value, rc

    def loglike_batch(self, thetas, nthreads=1):
        t = np.ascontiguousarray(np.atleast_2d(thetas), dtype=np.float32)
        w, stride = t.shape
        ll = np.zeros(w)
        st = np.zeros(w, dtype=np.int32)
        lib().orc_like2_loglike_batch(self.ptr, t.ctypes.data_as(FP), w,
                                      stride, stride, _dp(ll),
                                      st.ctypes.data_as(I32P), nthreads)
        return ll, st

    def initpd(self, b, theta, want_arrays=True):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        n = self.n
        rv = np.zeros(n * n) if want_arrays else None
        pr = np.zeros(2 * n * (n // 2 + 1)) if want_arrays else None
        sc = np.zeros(12)
        self._err(lib().orc_like2_initpd(
            self.ptr, b, t.ctypes.data_as(FP), t.size,
            _dp(rv) if rv is not None else C.cast(None, DP),
            _dp(pr) if pr is not None else C.cast(None, DP), _dp(sc)),
            "initpd")
        keys = ["dflux1", "dflux2", "mn1", "mn2", "var1", "var2", "shift1",
                "shift2", "doshift1", "doshift2", "r0", "wrap1"]
        out = dict(zip(keys, sc))
        if want_arrays:
            out["rvals"] = rv.reshape(n, n)
            c = pr[0::2] + 1j * pr[1::2]
            out["prtrans"] = c.reshape(n, n // 2 + 1)
        return out

    def getpd(self, b, f, sigmult1=1.0, sigmult2=1.0, setlog=True):
        n = self.n
        pd = np.zeros(n * n)
        info = np.zeros(8)
        self._err(lib().orc_like2_getpd(self.ptr, b, f, sigmult1, sigmult2,
                                        int(setlog), _dp(pd), _dp(info)),
                  "getpd")
        keys = ["minflux1", "minflux2", "dflux1", "dflux2", "split1",
                "split2", "sgpos1", "sgpos2"]
        return pd.reshape(n, n), dict(zip(keys, info))
