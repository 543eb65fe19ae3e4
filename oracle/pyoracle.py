"""ctypes front end of the CPU oracle (oracle/pofd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the reported CPU
baseline, never as the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "_build", "libpofd_oracle.so")
_lib = None

D = C.c_double
DP = C.POINTER(C.c_double)
U32P = C.POINTER(C.c_uint32)
I32P = C.POINTER(C.c_int32)
FP = C.POINTER(C.c_float)


class OrcErr(C.Structure):
    _fields_ = [("code", C.c_int), ("cls", C.c_char * 48),
                ("method", C.c_char * 48), ("msg", C.c_char * 512)]

    def text(self):
        return "%s::%s: %s" % (self.cls.decode(), self.method.decode(),
                               self.msg.decode())


class OracleError(RuntimeError):
    def __init__(self, code, text):
        super().__init__(text)
        self.code = code


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIBPATH):
        build()
    L = C.CDLL(_LIBPATH)
    sig = {
        "orc_cspline_init": (C.c_int, [C.c_size_t, DP, DP, DP]),
        "orc_cspline_eval": (D, [C.c_size_t, DP, DP, DP, D]),
        "orc_beam_new": (C.c_void_p, []),
        "orc_beam_free": (None, [C.c_void_p]),
        "orc_beam_set": (C.c_int, [C.c_void_p, DP, C.c_size_t, D, D,
                                   C.POINTER(OrcErr)]),
        "orc_beam_make_histogram": (C.c_int, [C.c_void_p, C.c_uint32,
                                              C.POINTER(OrcErr)]),
        "orc_beam_info": (None, [C.c_void_p, DP]),
        "orc_beam_array": (DP, [C.c_void_p, C.c_int]),
        "orc_knots_set_positions": (C.c_int, [C.c_void_p, C.c_uint, DP,
                                              C.POINTER(OrcErr)]),
        "orc_knots_set_params": (C.c_int, [C.c_void_p, C.c_uint, FP,
                                           C.POINTER(OrcErr)]),
        "orc_knots_is_valid": (C.c_int, [C.c_void_p]),
        "orc_knots_number_counts": (D, [C.c_void_p, D]),
        "orc_knots_rrange": (None, [C.c_void_p, C.c_void_p, DP, DP]),
        "orc_knots_getR1": (D, [C.c_void_p, D, C.c_void_p]),
        "orc_knots_getR": (None, [C.c_void_p, C.c_uint, DP, C.c_void_p, DP]),
        "orc_knots_regularize": (D, [C.c_void_p, D]),
        "orc_knots_splineint": (D, [C.c_void_p, D]),
        "orc_gl_nodes": (None, [C.c_uint, DP, DP]),
        "orc_data_bin": (C.c_int, [DP, C.c_size_t, C.c_uint32, U32P, DP, DP]),
        "orc_data_mean_subtract": (D, [DP, C.c_size_t]),
        "orc_rfft": (C.c_int, [C.c_uint, DP, DP]),
        "orc_irfft": (C.c_int, [C.c_uint, DP, DP]),
        "orc_pd_integral": (D, [C.c_uint, DP, D, C.c_int]),
        "orc_pd_mean": (D, [C.c_uint, DP, D, D, C.c_int, C.c_int]),
        "orc_pd_mean_var": (None, [C.c_uint, DP, D, D, C.c_int, C.c_int,
                                   DP, DP]),
        "orc_pd_interp": (D, [C.c_uint, DP, D, D, D]),
        "orc_pd_loglike_binned": (D, [C.c_uint, DP, D, D, C.c_int, U32P,
                                      C.c_uint32, D, D]),
        "orc_pd_loglike_unbinned": (D, [C.c_uint, DP, D, D, C.c_int, DP,
                                        C.c_size_t, DP]),
        "orc_like_new": (C.c_void_p, [C.c_uint, C.c_uint]),
        "orc_like_free": (None, [C.c_void_p]),
        "orc_like_error": (C.POINTER(OrcErr), [C.c_void_p]),
        "orc_like_set_precise": (None, [C.c_void_p, C.c_int]),
        "orc_like_set_split_shift": (None, [C.c_void_p, C.c_int]),
        "orc_like_set_timing": (None, [C.c_void_p, C.c_int]),
        "orc_like_timing": (None, [C.c_void_p, DP]),
        "orc_find_split_array": (C.c_int, [DP, C.c_uint, D, D, D, D, C.c_uint,
                                           C.POINTER(C.c_uint), DP,
                                           C.POINTER(OrcErr)]),
        "orc_like_set_knot_positions": (C.c_int, [C.c_void_p, C.c_uint, DP]),
        "orc_like_add_beamset": (C.c_int, [C.c_void_p, C.c_void_p]),
        "orc_like_add_field": (C.c_int, [C.c_void_p, C.c_int, DP, C.c_size_t,
                                         U32P, C.c_uint32, D, D, D, D]),
        "orc_like_finalize_setup": (C.c_int, [C.c_void_p, FP, C.c_uint, D]),
        "orc_like_beamset_info": (C.c_int, [C.c_void_p, C.c_int, DP]),
        "orc_like_loglike": (C.c_int, [C.c_void_p, FP, C.c_uint, DP]),
        "orc_like_loglike_batch": (C.c_int, [C.c_void_p, FP, C.c_uint,
                                             C.c_uint, C.c_uint, DP, I32P,
                                             C.c_uint]),
        "orc_like_initpd": (C.c_int, [C.c_void_p, C.c_int, FP, C.c_uint, DP,
                                      DP, DP, DP, DP]),
        "orc_like_getpd": (C.c_int, [C.c_void_p, C.c_int, C.c_int, D, C.c_int,
                                     DP, DP, DP]),
        "orc_like_rinterp_neg": (C.c_int, [C.c_void_p, DP, DP]),
        "orc_like_field_loglike": (C.c_int, [C.c_void_p, C.c_int, C.c_int,
                                             DP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(DP)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _nullable(a, ptype):
    return C.cast(None, ptype) if a is None else a.ctypes.data_as(ptype)


# --------------------------------------------------------------------------
def cspline_init(x, y):
    x, y = _f64(x), _f64(y)
    c = np.zeros_like(x)
    rc = lib().orc_cspline_init(len(x), _dp(x), _dp(y), _dp(c))
    if rc:
        raise OracleError(rc, "cspline_init failed")
    return c


def cspline_eval(x, y, c, xv):
    x, y, c = _f64(x), _f64(y), _f64(c)
    return np.array([lib().orc_cspline_eval(len(x), _dp(x), _dp(y), _dp(c),
                                            float(v))
                     for v in np.atleast_1d(xv)])


class Beam:
    """beam::readFile (+ makeHistogram) on an in-memory pixel image."""

    def __init__(self, pix, pixsize, minval=1e-5, nhist=None):
        self.ptr = lib().orc_beam_new()
        pix = _f64(np.asarray(pix).ravel())
        err = OrcErr()
        rc = lib().orc_beam_set(self.ptr, _dp(pix), pix.size, float(pixsize),
                                float(minval), C.byref(err))
        if rc:
            raise OracleError(rc, err.text())
        if nhist:
            rc = lib().orc_beam_make_histogram(self.ptr, int(nhist),
                                               C.byref(err))
            if rc:
                raise OracleError(rc, err.text())

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.orc_beam_free(self.ptr)
            self.ptr = None

    def info(self):
        out = np.zeros(16)
        lib().orc_beam_info(self.ptr, _dp(out))
        keys = ["npos", "nneg", "haspos", "hasneg", "pixsize", "minval",
                "is_pos_hist", "is_neg_hist", "posnbins", "negnbins",
                "area", "area_pos", "area_neg", "area_pix", "area_sq",
                "maxpos"]
        return dict(zip(keys, out))

    def array(self, which):
        names = ["pospix", "negpix", "posinv", "neginv", "posweights",
                 "poshistval", "negweights", "neghistval"]
        i = names.index(which)
        inf = self.info()
        n = {0: inf["npos"], 1: inf["nneg"], 2: inf["npos"], 3: inf["nneg"],
             4: inf["posnbins"], 5: inf["posnbins"], 6: inf["negnbins"],
             7: inf["negnbins"]}[i]
        p = lib().orc_beam_array(self.ptr, i)
        if not p or int(n) == 0:
            return np.zeros(0)
        return np.ctypeslib.as_array(p, shape=(int(n),)).copy()


class OrcKnots(C.Structure):
    _fields_ = [("nknots", C.c_uint), ("loaded", C.c_int),
                ("knots", C.c_double * 64), ("logknots", C.c_double * 64),
                ("logknotvals", C.c_double * 64), ("c", C.c_double * 64)]


class Knots:
    """numberCountsKnotsSpline: positions (double) and float32 log10 values."""

    def __init__(self, positions, values=None):
        self.s = OrcKnots()
        pos = _f64(positions)
        err = OrcErr()
        rc = lib().orc_knots_set_positions(C.byref(self.s), len(pos), _dp(pos),
                                           C.byref(err))
        if rc:
            raise OracleError(rc, err.text())
        if values is not None:
            self.set_params(values)

    def set_params(self, values):
        v = np.ascontiguousarray(values, dtype=np.float32)
        err = OrcErr()
        rc = lib().orc_knots_set_params(C.byref(self.s), len(v),
                                        v.ctypes.data_as(FP), C.byref(err))
        if rc:
            raise OracleError(rc, err.text())

    def is_valid(self):
        return bool(lib().orc_knots_is_valid(C.byref(self.s)))

    def number_counts(self, s):
        return np.array([lib().orc_knots_number_counts(C.byref(self.s),
                                                       float(x))
                         for x in np.atleast_1d(s)])

    def rrange(self, beam):
        a, b = D(), D()
        lib().orc_knots_rrange(C.byref(self.s), beam.ptr, C.byref(a),
                               C.byref(b))
        return a.value, b.value

    def getR1(self, f, beam):
        return np.array([lib().orc_knots_getR1(C.byref(self.s), float(x),
                                               beam.ptr)
                         for x in np.atleast_1d(f)])

    def getR(self, f, beam):
        f = _f64(np.atleast_1d(f))
        R = np.zeros_like(f)
        lib().orc_knots_getR(C.byref(self.s), len(f), _dp(f), beam.ptr, _dp(R))
        return R

    def regularize(self, alpha):
        return lib().orc_knots_regularize(C.byref(self.s), float(alpha))

    def splineint(self, alpha):
        """int S^alpha n(S) dS over the knot range (getNS / getFluxPerArea /
        getFluxSqPerArea for alpha = 0, 1, 2)."""
        return lib().orc_knots_splineint(C.byref(self.s), float(alpha))


def data_bin(d, nbins):
    d = _f64(d)
    bins = np.zeros(nbins, dtype=np.uint32)
    c0, dl = D(), D()
    rc = lib().orc_data_bin(_dp(d), d.size, nbins,
                            bins.ctypes.data_as(U32P), C.byref(c0),
                            C.byref(dl))
    if rc:
        raise OracleError(rc, "data_bin")
    return bins, c0.value, dl.value


def data_mean_subtract(d):
    d = _f64(d).copy()
    m = lib().orc_data_mean_subtract(_dp(d), d.size)
    return d, m


def rfft(x):
    x = _f64(x)
    n = x.size
    out = np.zeros(2 * (n // 2 + 1))
    rc = lib().orc_rfft(n, _dp(x), _dp(out))
    if rc:
        raise OracleError(rc, "rfft size")
    return out[0::2] + 1j * out[1::2]


def irfft(X, n):
    X = np.asarray(X, dtype=np.complex128)
    buf = np.zeros(2 * (n // 2 + 1))
    buf[0::2] = X.real
    buf[1::2] = X.imag
    out = np.zeros(n)
    rc = lib().orc_irfft(n, _dp(buf), _dp(out))
    if rc:
        raise OracleError(rc, "irfft size")
    return out


def pd_integral(pd, dflux, islog=False):
    pd = _f64(pd)
    return lib().orc_pd_integral(pd.size, _dp(pd), dflux, int(islog))


def pd_mean(pd, minflux, dflux, islog=False, donorm=True):
    pd = _f64(pd)
    return lib().orc_pd_mean(pd.size, _dp(pd), minflux, dflux, int(islog),
                             int(donorm))


def pd_mean_var(pd, minflux, dflux, islog=False, donorm=True):
    pd = _f64(pd)
    m, v = D(), D()
    lib().orc_pd_mean_var(pd.size, _dp(pd), minflux, dflux, int(islog),
                          int(donorm), C.byref(m), C.byref(v))
    return m.value, v.value


def pd_interp(pd, minflux, dflux, x):
    pd = _f64(pd)
    return np.array([lib().orc_pd_interp(pd.size, _dp(pd), minflux, dflux,
                                         float(v)) for v in np.atleast_1d(x)])


def pd_loglike_binned(pd, minflux, dflux, bins, bincent0, bindelta,
                      islog=True):
    pd = _f64(pd)
    bins = np.ascontiguousarray(bins, dtype=np.uint32)
    return lib().orc_pd_loglike_binned(pd.size, _dp(pd), minflux, dflux,
                                       int(islog), bins.ctypes.data_as(U32P),
                                       bins.size, bincent0, bindelta)


def pd_loglike_unbinned(pd, minflux, dflux, flux, islog=True):
    pd, flux = _f64(pd), _f64(flux)
    comp = D()
    v = lib().orc_pd_loglike_unbinned(pd.size, _dp(pd), minflux, dflux,
                                      int(islog), _dp(flux), flux.size,
                                      C.byref(comp))
    return v, comp.value


def find_split(pofd, dflux, sgpos, sgneg, minflux_r, wrap_ridx):
    """PDFactory::findSplitPoint (src/PDFactory.cc:691-921) on `pofd`.
    Returns (split index, peak); raises OracleError where it throws."""
    a = _f64(pofd)
    out = C.c_uint(0)
    peak = C.c_double(0.0)
    err = OrcErr()
    rc = lib().orc_find_split_array(_dp(a), a.size, dflux, sgpos, sgneg,
                                    minflux_r, wrap_ridx, C.byref(out),
                                    C.byref(peak), C.byref(err))
    if rc:
        raise OracleError(rc, err.text())
    return int(out.value), float(peak.value)


class Like:
    """calcLike restated: beam sets, fields, R ranges, getLogLike."""

    def __init__(self, fftsize, ninterp, knot_pos):
        self.ptr = lib().orc_like_new(fftsize, ninterp)
        self.fftsize = fftsize
        self.ninterp = ninterp
        self.nknots = len(knot_pos)
        kp = _f64(knot_pos)
        rc = lib().orc_like_set_knot_positions(self.ptr, len(kp), _dp(kp))
        self._check(rc)

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.orc_like_free(self.ptr)
            self.ptr = None

    def _check(self, rc):
        if rc:
            e = lib().orc_like_error(self.ptr).contents
            raise OracleError(rc, e.text())

    def set_precise(self, precise=True):
        """1 (True): extended-precision transforms; 0: double (reference
        arithmetic); 2 and 3: analysis modes (extended forward + double
        inverse; extended transforms with X_k, R^_0 rounded to double)."""
        lib().orc_like_set_precise(self.ptr, int(precise))

    TIME_KEYS = ("R", "p0", "fft", "pos", "split", "copy", "norm", "mean",
                 "log", "like", "other")

    def set_timing(self, on=True):
        """Per-stage timing of the batch workers in the reference's TIMING
        categories (src/PDFactory.cc:93-125) plus the likelihood sum and
        the rest; resets the totals."""
        lib().orc_like_set_timing(self.ptr, int(on))

    def timing(self):
        """Thread CPU seconds per stage since the last set_timing."""
        out = np.zeros(len(self.TIME_KEYS))
        lib().orc_like_timing(self.ptr, _dp(out))
        return dict(zip(self.TIME_KEYS, out.tolist()))

    def set_split_shift(self, shift):
        """Test hook: move every split index found by `shift` before the
        rotation (0 restores the reference's behaviour)."""
        lib().orc_like_set_split_shift(self.ptr, int(shift))

    def add_beamset(self, beam):
        return lib().orc_like_add_beamset(self.ptr, beam.ptr)

    def add_field(self, b, pix=None, bins=None, bincent0=0.0, bindelta=0.0,
                  sigma_base=0.0025, like_norm=1.0, npix=None):
        p = None if pix is None else _f64(pix)
        bb = None if bins is None else np.ascontiguousarray(bins, np.uint32)
        n = npix if npix is not None else (p.size if p is not None else 0)
        return lib().orc_like_add_field(
            self.ptr, b, _nullable(p, DP), n, _nullable(bb, U32P),
            0 if bb is None else bb.size, bincent0, bindelta, sigma_base,
            like_norm)

    def finalize(self, theta, exp_conf=0.0):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        rc = lib().orc_like_finalize_setup(self.ptr, t.ctypes.data_as(FP),
                                           t.size, exp_conf)
        self._check(rc)

    def beamset_info(self, b):
        out = np.zeros(8)
        self._check(lib().orc_like_beamset_info(self.ptr, b, _dp(out)))
        return dict(zip(["minRFlux", "maxRFlux", "maxsigma_base",
                         "minDataFlux", "maxDataFlux", "nfields", "exp_conf"],
                        out))

    def loglike(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        v = D()
        rc = lib().orc_like_loglike(self.ptr, t.ctypes.data_as(FP), t.size,
                                    C.byref(v))
        self._check(rc)
        return v.value

    def loglike_batch(self, thetas, nthreads=1):
        t = np.ascontiguousarray(thetas, dtype=np.float32)
        w, npar = t.shape
        out = np.zeros(w)
        st = np.zeros(w, dtype=np.int32)
        lib().orc_like_loglike_batch(self.ptr, t.ctypes.data_as(FP), w, npar,
                                     npar, _dp(out), st.ctypes.data_as(I32P),
                                     nthreads)
        return out, st

    def initpd(self, b, theta):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        n = self.fftsize
        rf = np.zeros(self.ninterp)
        rv = np.zeros(self.ninterp)
        rvals = np.zeros(n)
        pr = np.zeros(2 * (n // 2 + 1))
        sc = np.zeros(8)
        rc = lib().orc_like_initpd(self.ptr, b, t.ctypes.data_as(FP), t.size,
                                   _dp(rf), _dp(rv), _dp(rvals), _dp(pr),
                                   _dp(sc))
        self._check(rc)
        return dict(rinterp_flux=rf, rinterp_vals=rv, rvals=rvals,
                    prtrans=pr[0::2] + 1j * pr[1::2], dflux=sc[0],
                    minflux_R=sc[1], wrapRidx=int(sc[2]), mn=sc[3],
                    var_noi=sc[4], doshift=bool(sc[5]), shift=sc[6], r0=sc[7])

    def rinterp_neg(self):
        """Negative-beam interpolation nodes and log2 R of the last initpd."""
        rf = np.zeros(self.ninterp)
        rv = np.zeros(self.ninterp)
        self._check(lib().orc_like_rinterp_neg(self.ptr, _dp(rf), _dp(rv)))
        return rf, rv

    def getpd(self, b, f, sigmult=1.0, setlog=True):
        n = self.fftsize
        raw = np.zeros(n)
        pd = np.zeros(n)
        info = np.zeros(8)
        rc = lib().orc_like_getpd(self.ptr, b, f, sigmult, int(setlog),
                                  _dp(raw), _dp(pd), _dp(info))
        res = dict(pofd_raw=raw, pd=pd, minflux=info[0], dflux=info[1],
                   split=int(info[2]), peak=info[3], sgpos=info[4],
                   sgneg=info[5], integral=info[6], islog=bool(info[7]))
        if rc:
            e = lib().orc_like_error(self.ptr).contents
            err = OracleError(rc, e.text())
            err.partial = res
            raise err
        return res

    def field_loglike(self, b, f):
        out = np.zeros(2)
        self._check(lib().orc_like_field_loglike(self.ptr, b, f, _dp(out)))
        return out[0], out[1]
