"""ctypes binding of libpofd_affine_amd.so (the C ABI in include/).

The library is built in-tree by ``pofd_affine_amd/csrc/Makefile`` (see
``__graft_entry__.build()``).  There is no CPU fallback: if the library or a
gfx950 device is missing, the calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "_build", "libpofd_affine_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "pofd_affine_amd.h")

POFD_OK = 0
POFD_E_SPLIT = 1
POFD_E_INVALID_MODEL = 2
POFD_E_MEAN = 3
POFD_E_BADLIKE = 4
POFD_E_ARG = 5
POFD_E_INTERNAL = 6
POFD_E_NOGPU = 7
POFD_OPT_SPECTRAL_CUT = 1
POFD_OPT_STAGE_TIMING = 2
STAGES = ("prep", "forward", "inverse", "finalize", "pixels", "combine")

STATUS_NAMES = {0: "ok", 1: "split", 2: "invalid_model", 3: "mean",
                4: "bad_like", 5: "arg", 6: "internal", 7: "nogpu"}

D = C.c_double
DP = C.POINTER(C.c_double)
U32P = C.POINTER(C.c_uint32)
I32P = C.POINTER(C.c_int32)
FP = C.POINTER(C.c_float)
VP = C.c_void_p

_SIGS = {
    "pofd_create": (C.c_int, [C.POINTER(VP), C.c_int]),
    "pofd_destroy": (None, [VP]),
    "pofd_last_error": (C.c_char_p, [VP]),
    "pofd_set_option": (C.c_int, [VP, C.c_int, D]),
    "pofd_beam_create": (C.c_int, [C.POINTER(VP), DP, C.c_uint64, D, D]),
    "pofd_beam_make_histogram": (C.c_int, [VP, C.c_uint32]),
    "pofd_beam_destroy": (None, [VP]),
    "pofd_beam_get": (C.c_int, [VP, C.c_int, DP, U32P]),
    "pofd_beam_info": (C.c_int, [VP, DP]),
    "pofd_data_bin": (C.c_int, [DP, C.c_uint64, C.c_uint32, U32P, DP, DP]),
    "pofd_data_mean_subtract": (C.c_int, [DP, C.c_uint64, DP]),
    "pofd1d_configure": (C.c_int, [VP, C.c_uint32, C.c_uint32, DP,
                                   C.c_uint32, D]),
    "pofd1d_add_beamset": (C.c_int, [VP, VP, C.POINTER(C.c_int)]),
    "pofd1d_add_field": (C.c_int, [VP, C.c_int, DP, C.c_uint64, U32P,
                                   C.c_uint32, D, D, D, D,
                                   C.POINTER(C.c_int)]),
    "pofd1d_prepare": (C.c_int, [VP]),
    "pofd1d_set_parts": (C.c_int, [VP, C.c_int]),
    "pofd1d_autotune_parts": (C.c_int, [VP, FP, C.c_uint32, I32P, C.c_uint32,
                                        C.c_uint32, I32P]),
    "pofd1d_graph_create": (C.c_int, [C.POINTER(VP), C.c_uint32, VP,
                                      C.c_uint32, C.c_uint32, VP, VP, VP,
                                      C.POINTER(VP)]),
    "pofd1d_graph_launch": (C.c_int, [VP, VP]),
    "pofd1d_graph_destroy": (None, [VP]),
    "pofd1d_set_priors": (C.c_int, [VP, C.c_int, D, D, D, D, D, D]),
    "pofd1d_loglike_batch_device_count": (C.c_int, [VP, VP, C.c_uint32, VP,
                                                    C.c_uint32, VP, VP, VP]),
    "pofd1d_set_rrange": (C.c_int, [VP, C.c_int, D, D]),
    "pofd_pd_loglike": (C.c_int, [VP, DP, C.c_uint32, D, D, C.c_int, DP,
                                  C.c_uint64, U32P, C.c_uint32, D, D, DP]),
    "pofd_pd2_loglike": (C.c_int, [VP, DP, C.c_uint32, C.c_uint32, D, D, D,
                                   D, C.c_int, DP, DP, C.c_uint64, U32P,
                                   C.c_uint32, C.c_uint32, D, D, D, D, DP]),
    "pofd1d_beamset_info": (C.c_int, [VP, C.c_int, DP]),
    "pofd1d_loglike_batch": (C.c_int, [VP, FP, C.c_uint32, C.c_uint32, DP,
                                       I32P]),
    "pofd1d_loglike_batch_device": (C.c_int, [VP, VP, C.c_uint32,
                                              C.c_uint32, VP, VP, VP]),
    "pofd1d_loglike_batch_device_noalloc": (C.c_int, [VP, VP, C.c_uint32,
                                                      C.c_uint32, VP, VP,
                                                      VP]),
    "pofd1d_get_pd": (C.c_int, [VP, C.c_uint32, C.c_int, C.c_int, C.c_int,
                                DP, DP, DP]),
    "pofd1d_get_debug": (C.c_int, [VP, C.c_uint32, C.c_int, C.c_int, DP]),
    "pofd1d_get_spectrum": (C.c_int, [VP, C.c_uint32, C.c_int, DP,
                                      C.c_uint32, U32P]),
    "pofd1d_get_rinterp": (C.c_int, [VP, C.c_uint32, C.c_int, DP, DP]),
    "pofd1d_get_rinterp_neg": (C.c_int, [VP, C.c_uint32, C.c_int, DP, DP]),
    "pofd1d_get_bonus": (C.c_int, [VP, C.c_uint32, DP]),
    "pofd2d_get_bonus": (C.c_int, [VP, C.c_uint32, DP]),
    "pofd_ens_create": (C.c_int, [C.POINTER(VP), VP, C.c_int, C.c_uint32,
                                  C.c_uint32, D, C.c_uint64]),
    "pofd_ens_destroy": (None, [VP]),
    "pofd_ens_last_error": (C.c_char_p, [VP]),
    "pofd_ens_set_param_state": (C.c_int, [VP, C.POINTER(C.c_uint8)]),
    "pofd_ens_set_limits": (C.c_int, [VP, FP, FP]),
    "pofd_ens_set_exponent_mode": (C.c_int, [VP, C.c_int, FP]),
    "pofd_ens_set_gauss_target": (C.c_int, [VP, DP, DP]),
    "pofd_ens_set_positions": (C.c_int, [VP, FP]),
    "pofd_ens_run": (C.c_int, [VP, C.c_uint32, D, C.c_uint32]),
    "pofd_ens_get_state": (C.c_int, [VP, FP, DP, U32P, U32P]),
    "pofd_ens_chain_length": (C.c_int, [VP, U32P]),
    "pofd_ens_get_chain": (C.c_int, [VP, C.c_uint32, C.c_uint32, FP, DP]),
    "pofd_ens_clear_chain": (C.c_int, [VP]),
    "pofd_ens_acor": (C.c_int, [VP, C.c_uint32, DP, DP, DP]),
    "pofd_ens_sample_z": (C.c_int, [VP, C.c_uint32, DP]),
    "pofd_acor": (C.c_int, [FP, C.c_uint32, DP, DP, DP]),
    "pofd_ens_set_surrogate": (C.c_int, [VP, VP, U32P, C.c_uint32, D, D,
                                         C.POINTER(C.c_int)]),
    "pofd_ens_set_surrogate_steps": (C.c_int, [VP, C.c_uint32]),
    "pofd_ens_counts": (C.c_int, [VP, C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_uint64)]),
    "pofd_mlp_create": (C.c_int, [C.POINTER(VP), C.c_uint32, C.c_uint32,
                                  C.c_uint64]),
    "pofd_mlp_destroy": (None, [VP]),
    "pofd_mlp_last_error": (C.c_char_p, [VP]),
    "pofd_mlp_info": (C.c_int, [VP, U32P, U32P]),
    "pofd_mlp_nparams": (C.c_int, [VP, U32P]),
    "pofd_mlp_layout": (C.c_int, [VP, C.POINTER(C.c_uint64)]),
    "pofd_mlp_get_params": (C.c_int, [VP, FP]),
    "pofd_mlp_set_params": (C.c_int, [VP, FP]),
    "pofd_mlp_set_standardization": (C.c_int, [VP, FP, FP, D, D]),
    "pofd_mlp_set_whitening": (C.c_int, [VP, FP, FP, D, D]),
    "pofd_mlp_predict": (C.c_int, [VP, FP, C.c_uint32, DP]),
    "pofd_mlp_loss_grad": (C.c_int, [VP, FP, DP, C.c_uint32, FP, DP]),
    "pofd_mlp_train": (C.c_int, [VP, FP, DP, C.c_uint32, C.c_uint32,
                                 C.c_uint32, C.c_float, C.c_float, C.c_uint64,
                                 DP, DP]),
    "pofd1d_stage_times": (C.c_int, [VP, DP, U32P]),
    "pofd_dbeam_create": (C.c_int, [C.POINTER(VP), DP, DP, C.c_uint64, D, D]),
    "pofd_dbeam_make_histogram": (C.c_int, [VP, C.c_uint32]),
    "pofd_dbeam_destroy": (None, [VP]),
    "pofd_dbeam_last_error": (C.c_char_p, []),
    "pofd_dbeam_info": (C.c_int, [VP, DP]),
    "pofd_dbeam_get": (C.c_int, [VP, C.c_int, C.c_int, DP, U32P]),
    "pofd_data2_bin": (C.c_int, [DP, DP, C.c_uint64, C.c_uint32, C.c_uint32,
                                 U32P, DP]),
    "pofd_data2_mean_subtract": (C.c_int, [DP, DP, C.c_uint64, DP]),
    "pofd2d_configure": (C.c_int, [VP, C.c_uint32, C.c_uint32, C.c_int, DP,
                                   C.c_uint32, DP, C.c_uint32, DP,
                                   C.c_uint32, D, D, D]),
    "pofd2d_add_beamset": (C.c_int, [VP, VP, C.POINTER(C.c_int)]),
    "pofd2d_add_field": (C.c_int, [VP, C.c_int, DP, DP, C.c_uint64, U32P,
                                   C.c_uint32, C.c_uint32, DP, D, D, D,
                                   C.POINTER(C.c_int)]),
    "pofd2d_prepare": (C.c_int, [VP, FP, C.c_uint32]),
    "pofd2d_set_rrange": (C.c_int, [VP, C.c_int, DP]),
    "pofd2d_beamset_info": (C.c_int, [VP, C.c_int, DP]),
    "pofd2d_loglike_batch": (C.c_int, [VP, FP, C.c_uint32, C.c_uint32, DP,
                                       I32P]),
    "pofd2d_loglike_batch_device": (C.c_int, [VP, VP, C.c_uint32,
                                              C.c_uint32, VP, VP, VP]),
    "pofd2d_get_debug": (C.c_int, [VP, C.c_uint32, C.c_int, C.c_int, DP]),
    "pofd2d_get_pd": (C.c_int, [VP, C.c_uint32, C.c_int, C.c_int, C.c_int,
                                DP, DP]),
}

_lib = None


class PofdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (status %d: %s)" % (msg, code,
                                                 STATUS_NAMES.get(code, "?")))
        self.code = code


def exported_symbols():
    """Names the public header declares (parsed from include/)."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(pofd(?:1d)?_[a-z_]+)\s*\(", txt)))


def lib():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise OSError("libpofd_affine_amd.so is not built (%s); run "
                      "__graft_entry__.build()" % LIBPATH)
    L = C.CDLL(LIBPATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(DP)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- host prep
class Beam:
    """beam::readFile + makeHistogram restated (host, bit-exact)."""

    def __init__(self, pix, pixsize, minval=1e-5, nhist=None):
        pix = _f64(np.asarray(pix).ravel())
        self.ptr = VP()
        rc = lib().pofd_beam_create(C.byref(self.ptr), _dp(pix), pix.size,
                                    float(pixsize), float(minval))
        if rc:
            raise PofdError(rc, "beam::readFile failed")
        if nhist:
            rc = lib().pofd_beam_make_histogram(self.ptr, int(nhist))
            if rc:
                raise PofdError(rc, "beam::makeHistogram failed")

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.pofd_beam_destroy(self.ptr)
            self.ptr = None

    def info(self):
        out = np.zeros(16)
        lib().pofd_beam_info(self.ptr, _dp(out))
        keys = ["npos", "nneg", "haspos", "hasneg", "pixsize", "minval",
                "is_pos_hist", "is_neg_hist", "posnbins", "negnbins",
                "area", "area_pos", "area_neg", "area_pix", "area_sq",
                "maxpos"]
        return dict(zip(keys, out))

    def array(self, which):
        names = ["pospix", "negpix", "posinv", "neginv", "posweights",
                 "poshistval", "negweights", "neghistval"]
        i = names.index(which)
        n = C.c_uint32()
        lib().pofd_beam_get(self.ptr, i, None, C.byref(n))
        out = np.zeros(n.value)
        if n.value:
            lib().pofd_beam_get(self.ptr, i, _dp(out), C.byref(n))
        return out


def data_bin(d, nbins):
    """fitsData::applyBinning restated (host, bit-exact)."""
    d = _f64(d)
    bins = np.zeros(nbins, dtype=np.uint32)
    c0, dl = D(), D()
    rc = lib().pofd_data_bin(_dp(d), d.size, nbins,
                             bins.ctypes.data_as(U32P), C.byref(c0),
                             C.byref(dl))
    if rc:
        raise PofdError(rc, "fitsData::applyBinning failed")
    return bins, c0.value, dl.value


def data_mean_subtract(d):
    d = _f64(d).copy()
    m = D()
    rc = lib().pofd_data_mean_subtract(_dp(d), d.size, C.byref(m))
    if rc:
        raise PofdError(rc, "fitsData::meanSubtract failed")
    return d, m.value


# ------------------------------------------------------------------ context
class Context:
    """One pofd_ctx: a 1D likelihood (calcLike) on one gfx950 device."""

    def __init__(self, device=0):
        self.ptr = VP()
        rc = lib().pofd_create(C.byref(self.ptr), int(device))
        if rc:
            raise PofdError(rc, "pofd_create failed (no gfx950 device?)")
        self._keep = []

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.pofd_destroy(self.ptr)
            self.ptr = None

    def _check(self, rc, what):
        if rc:
            msg = lib().pofd_last_error(self.ptr)
            raise PofdError(rc, "%s: %s" % (what, msg.decode() if msg else ""))

    def set_spectral_cut(self, value):
        self._check(lib().pofd_set_option(self.ptr, POFD_OPT_SPECTRAL_CUT,
                                          float(value)), "pofd_set_option")

    def set_priors(self, sigma_width=None, cfirb=None, poisson=None,
                   reg_alpha=0.0):
        """Priors added on the device to every batch (calcLike::getLogLike):
        sigma_width, cfirb=(mean, sigma), poisson=(mean, sigma)."""
        flags = 0
        sw = cm = cs = pm = ps = 0.0
        if sigma_width is not None:
            flags |= 1
            sw = sigma_width
        if cfirb is not None:
            flags |= 2
            cm, cs = cfirb
        if poisson is not None:
            flags |= 4
            pm, ps = poisson
        self._check(lib().pofd1d_set_priors(self.ptr, flags, sw, cm, cs, pm,
                                            ps, reg_alpha),
                    "pofd1d_set_priors")

    def set_parts(self, parts):
        """Workgroups per walker of the 1D transforms (0: by batch size)."""
        self._check(lib().pofd1d_set_parts(self.ptr, int(parts)),
                    "pofd1d_set_parts")

    def autotune_parts(self, thetas, sizes, reps=5):
        """Measured workgroups per walker for each batch size in `sizes`
        (pofd1d_autotune_parts); returns the chosen split per size."""
        th = np.ascontiguousarray(thetas, dtype=np.float32)
        sz = np.ascontiguousarray(sizes, dtype=np.int32)
        out = np.zeros(len(sz), dtype=np.int32)
        self._check(lib().pofd1d_autotune_parts(
            self.ptr, th.ctypes.data_as(FP), th.shape[1], sz.ctypes.data_as(I32P),
            len(sz), int(reps), out.ctypes.data_as(I32P)), "pofd1d_autotune_parts")
        return out

    def set_stage_timing(self, on):
        self._check(lib().pofd_set_option(self.ptr, POFD_OPT_STAGE_TIMING,
                                          1.0 if on else 0.0),
                    "pofd_set_option")

    def stage_times(self):
        """{stage: summed GPU ms} over the launches since the last call, and
        the number of beam-set launches counted."""
        out = np.zeros(len(STAGES))
        n = C.c_uint32()
        self._check(lib().pofd1d_stage_times(self.ptr, _dp(out), C.byref(n)),
                    "pofd1d_stage_times")
        return dict(zip(STAGES, out)), n.value

    def configure(self, nflux, ninterp, knot_pos, exp_conf=0.0):
        kp = _f64(knot_pos)
        self.nknots = kp.size
        self.nflux = nflux
        self.ninterp = ninterp
        self._check(lib().pofd1d_configure(self.ptr, nflux, ninterp, _dp(kp),
                                           kp.size, float(exp_conf)),
                    "pofd1d_configure")

    def add_beamset(self, beam):
        i = C.c_int()
        self._check(lib().pofd1d_add_beamset(self.ptr, beam.ptr, C.byref(i)),
                    "pofd1d_add_beamset")
        return i.value

    def add_field(self, b, pix=None, bins=None, bincent0=0.0, bindelta=0.0,
                  sigma_base=0.0025, like_norm=1.0, npix=None):
        i = C.c_int()
        if bins is not None:
            bb = np.ascontiguousarray(bins, dtype=np.uint32)
            n = int(npix if npix is not None else bb.sum())
            rc = lib().pofd1d_add_field(self.ptr, b, None, n,
                                        bb.ctypes.data_as(U32P), bb.size,
                                        bincent0, bindelta, sigma_base,
                                        like_norm, C.byref(i))
        else:
            p = _f64(pix)
            rc = lib().pofd1d_add_field(self.ptr, b, _dp(p), p.size, None, 0,
                                        0.0, 0.0, sigma_base, like_norm,
                                        C.byref(i))
        self._check(rc, "pofd1d_add_field")
        return i.value

    def prepare(self):
        self._check(lib().pofd1d_prepare(self.ptr), "pofd1d_prepare")

    def set_rrange(self, b, mn, mx):
        self._check(lib().pofd1d_set_rrange(self.ptr, b, mn, mx),
                    "pofd1d_set_rrange")

    def beamset_info(self, b):
        out = np.zeros(8)
        self._check(lib().pofd1d_beamset_info(self.ptr, b, _dp(out)),
                    "pofd1d_beamset_info")
        return dict(zip(["minRFlux", "maxRFlux", "maxsigma_base",
                         "minDataFlux", "maxDataFlux", "nfields", "exp_conf",
                         "dflux"], out))

    def loglike_batch(self, thetas):
        t = np.ascontiguousarray(np.atleast_2d(thetas), dtype=np.float32)
        w, stride = t.shape
        ll = np.zeros(w)
        st = np.zeros(w, dtype=np.int32)
        self._check(lib().pofd1d_loglike_batch(self.ptr, t.ctypes.data_as(FP),
                                               w, stride, _dp(ll),
                                               st.ctypes.data_as(I32P)),
                    "pofd1d_loglike_batch")
        return ll, st

    def loglike_batch_device(self, d_theta, nwalk, stride, d_ll, d_st,
                             stream=0):
        """Device pointers (ints) on this context's device; asynchronous."""
        self._check(lib().pofd1d_loglike_batch_device(
            self.ptr, VP(d_theta), nwalk, stride, VP(d_ll), VP(d_st),
            VP(stream)), "pofd1d_loglike_batch_device")

    def loglike_batch_device_noalloc(self, d_theta, nwalk, stride, d_ll,
                                     d_st, stream=0):
        """loglike_batch_device with the workspace in the library's fixed
        device scratch: allocates nothing."""
        self._check(lib().pofd1d_loglike_batch_device_noalloc(
            self.ptr, VP(d_theta), nwalk, stride, VP(d_ll), VP(d_st),
            VP(stream)), "pofd1d_loglike_batch_device_noalloc")

    def loglike_batch_device_count(self, d_theta, max_walkers, d_count,
                                   stride, d_ll, d_st, stream=0):
        """As loglike_batch_device for the first *d_count rows, the count
        an int32 in device memory."""
        self._check(lib().pofd1d_loglike_batch_device_count(
            self.ptr, VP(d_theta), max_walkers, VP(d_count), stride,
            VP(d_ll), VP(d_st), VP(stream)),
            "pofd1d_loglike_batch_device_count")

    def pd_loglike(self, pd, minflux, dflux, is_log=True, pix=None,
                   bins=None, bincent0=0.0, bindelta=0.0):
        """PD::getLogLike of pixels (or a histogram) given a host P(D)."""
        a = _f64(pd)
        out = C.c_double(0.0)
        if bins is not None:
            bb = np.ascontiguousarray(bins, dtype=np.uint32)
            rc = lib().pofd_pd_loglike(self.ptr, _dp(a), a.size, minflux,
                                       dflux, int(is_log), None, 0,
                                       bb.ctypes.data_as(U32P), bb.size,
                                       bincent0, bindelta, C.byref(out))
        else:
            p = _f64(pix)
            rc = lib().pofd_pd_loglike(self.ptr, _dp(a), a.size, minflux,
                                       dflux, int(is_log), _dp(p), p.size,
                                       None, 0, 0.0, 0.0, C.byref(out))
        self._check(rc, "pofd_pd_loglike")
        return out.value

    def pd2_loglike(self, pd, minflux1, dflux1, minflux2, dflux2,
                    is_log=True, d1=None, d2=None, bins=None, bininfo=None):
        """PDDouble::getLogLike of pixel pairs (or a 2D histogram, bininfo =
        (bincent01, bindelta1, bincent02, bindelta2)) given a host P(D)."""
        a = _f64(pd)
        n1, n2 = a.shape
        out = C.c_double(0.0)
        if bins is not None:
            bb = np.ascontiguousarray(bins, dtype=np.uint32)
            nb1, nb2 = bb.shape
            info = [float(v) for v in bininfo]
            rc = lib().pofd_pd2_loglike(self.ptr, _dp(a), n1, n2, minflux1,
                                        dflux1, minflux2, dflux2, int(is_log),
                                        None, None, 0,
                                        bb.ctypes.data_as(U32P), nb1, nb2,
                                        *info, C.byref(out))
        else:
            p1, p2 = _f64(d1), _f64(d2)
            rc = lib().pofd_pd2_loglike(self.ptr, _dp(a), n1, n2, minflux1,
                                        dflux1, minflux2, dflux2, int(is_log),
                                        _dp(p1), _dp(p2), p1.size, None, 0, 0,
                                        0.0, 0.0, 0.0, 0.0, C.byref(out))
        self._check(rc, "pofd_pd2_loglike")
        return out.value

    def get_pd(self, walker, b=0, f=0, is_log=True):
        """is_log True/False: PD::pd_ after getPD; "raw": the clamped,
        unrotated, unnormalised transform output."""
        is_log = -1 if is_log == "raw" else int(bool(is_log))
        out = np.zeros(self.nflux)
        mf, df = D(), D()
        self._check(lib().pofd1d_get_pd(self.ptr, walker, b, f, is_log,
                                        _dp(out), C.byref(mf), C.byref(df)),
                    "pofd1d_get_pd")
        return out, mf.value, df.value

    DEBUG_KEYS = ["split", "peak", "minidx", "minval", "jitter", "integral",
                  "tmn", "sgpos", "sgneg", "mn", "var_noi", "r0", "kmax",
                  "nrows", "field_loglike", "status", "itot", "minflux",
                  "pofd_is_log", "kmax_beamset", "S", "SJ", "Shigh", "sigma",
                  "t_prep", "t_reserved", "t_forward", "t_spectral",
                  "t_inverse", "t_split", "t_loglike", "t_total"]

    def get_debug(self, walker, b=0, f=0):
        out = np.zeros(32)
        self._check(lib().pofd1d_get_debug(self.ptr, walker, b, f, _dp(out)),
                    "pofd1d_get_debug")
        return dict(zip(self.DEBUG_KEYS, out))

    def get_spectrum(self, walker, b=0):
        n = C.c_uint32()
        self._check(lib().pofd1d_get_spectrum(self.ptr, walker, b, None, 0,
                                              C.byref(n)), "spectrum size")
        out = np.zeros(2 * n.value)
        self._check(lib().pofd1d_get_spectrum(self.ptr, walker, b, _dp(out),
                                              n.value, C.byref(n)),
                    "pofd1d_get_spectrum")
        return out[0::2] + 1j * out[1::2]

    def get_rinterp(self, walker, b=0):
        x = np.zeros(self.ninterp)
        y = np.zeros(self.ninterp)
        self._check(lib().pofd1d_get_rinterp(self.ptr, walker, b, _dp(x),
                                             _dp(y)), "pofd1d_get_rinterp")
        return x, y

    def get_bonus(self, nwalk):
        """[nwalk, 3]: getNS, getFluxPerArea, getFluxSqPerArea per walker of
        the last batch."""
        out = np.zeros(3 * nwalk)
        self._check(lib().pofd1d_get_bonus(self.ptr, nwalk, _dp(out)),
                    "pofd1d_get_bonus")
        return out.reshape(nwalk, 3)

    def get_rinterp_neg(self, walker, b=0):
        """Negative-beam interpolation nodes (ascending, < 0) and log2 R."""
        x = np.zeros(self.ninterp)
        y = np.zeros(self.ninterp)
        self._check(lib().pofd1d_get_rinterp_neg(self.ptr, walker, b, _dp(x),
                                                 _dp(y)),
                    "pofd1d_get_rinterp_neg")
        return x, y


# ------------------------------------------------------------- two bands
class DoubleBeam:
    """doublebeam::setBeams + makeHistogram restated (host, bit-exact)."""

    KEYS = ["hassign", "npix", "ishist", "nhist", "tot1", "tot2", "totsq1",
            "totsq2", "min1", "max1", "min2", "max2"]

    def __init__(self, b1, b2, pixsize, minval=1e-6, nhist=None):
        b1, b2 = _f64(np.ravel(b1)), _f64(np.ravel(b2))
        assert b1.size == b2.size
        self.ptr = VP()
        rc = lib().pofd_dbeam_create(C.byref(self.ptr), _dp(b1), _dp(b2),
                                     b1.size, float(pixsize), float(minval))
        if rc:
            raise PofdError(rc, lib().pofd_dbeam_last_error().decode())
        if nhist:
            rc = lib().pofd_dbeam_make_histogram(self.ptr, int(nhist))
            if rc:
                raise PofdError(rc, lib().pofd_dbeam_last_error().decode())

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.pofd_dbeam_destroy(self.ptr)
            self.ptr = None

    def info(self):
        out = np.zeros(64)
        lib().pofd_dbeam_info(self.ptr, _dp(out))
        sgn = [dict(zip(self.KEYS, out[12 * k:12 * k + 12])) for k in range(4)]
        return sgn, dict(pixsize=out[48], minval=out[49], nbins=out[50])

    def array(self, sgn, which):
        names = ["pix1", "pix2", "ipix1", "ipix2", "logratio", "binw",
                 "binv1", "binv2", "binlogratio"]
        n = C.c_uint32()
        lib().pofd_dbeam_get(self.ptr, sgn, names.index(which), None,
                             C.byref(n))
        out = np.zeros(n.value)
        if n.value:
            lib().pofd_dbeam_get(self.ptr, sgn, names.index(which), _dp(out),
                                 C.byref(n))
        return out


def data2_bin(d1, d2, nb1, nb2):
    """fitsDataDouble::applyBinning restated (host, bit-exact)."""
    d1, d2 = _f64(d1), _f64(d2)
    bins = np.zeros(nb1 * nb2, dtype=np.uint32)
    info = np.zeros(4)
    rc = lib().pofd_data2_bin(_dp(d1), _dp(d2), d1.size, nb1, nb2,
                              bins.ctypes.data_as(U32P), _dp(info))
    if rc:
        raise PofdError(rc, "fitsDataDouble::applyBinning failed")
    return bins.reshape(nb1, nb2), info


def data2_mean_subtract(d1, d2):
    d1, d2 = _f64(d1).copy(), _f64(d2).copy()
    m = np.zeros(2)
    rc = lib().pofd_data2_mean_subtract(_dp(d1), _dp(d2), d1.size, _dp(m))
    if rc:
        raise PofdError(rc, "fitsDataDouble::meanSubtract failed")
    return d1, d2, m


class Context2(Context):
    """A pofd_ctx used for the two-band likelihood (calcLikeDouble)."""

    DEBUG_KEYS = ["status", "split1", "split2", "sg1pos", "sg2pos", "sigma1",
                  "sigma2", "integral", "itot", "minflux1", "minflux2",
                  "field_loglike", "mn1", "mn2", "var1", "var2", "r0", "K1",
                  "K2", "split_status1", "split_status2"]

    def configure2d(self, n, knots, sigmapos, offsetpos, nedge=256,
                    edge_integ=True, min_sigma=0.0, exp_conf1=0.0,
                    exp_conf2=0.0):
        kp, sp, op = _f64(knots), _f64(sigmapos), _f64(offsetpos)
        self.n = n
        self.nmodel = kp.size + sp.size + op.size
        self._check(lib().pofd2d_configure(self.ptr, n, nedge,
                                           int(edge_integ), _dp(kp), kp.size,
                                           _dp(sp), sp.size, _dp(op), op.size,
                                           min_sigma, exp_conf1, exp_conf2),
                    "pofd2d_configure")

    def add_beamset2d(self, dbeam):
        i = C.c_int()
        self._keep.append(dbeam)
        self._check(lib().pofd2d_add_beamset(self.ptr, dbeam.ptr, C.byref(i)),
                    "pofd2d_add_beamset")
        return i.value

    def add_field2d(self, b, d1, d2, sigma_base1, sigma_base2, like_norm=1.0,
                    bins=None, bininfo=None):
        d1, d2 = _f64(d1), _f64(d2)
        self._keep += [d1, d2]
        i = C.c_int()
        if bins is not None:
            bins = np.ascontiguousarray(bins, dtype=np.uint32)
            nb1, nb2 = bins.shape
            info = _f64(bininfo)
            self._keep += [bins, info]
            bp, ip = bins.ctypes.data_as(U32P), _dp(info)
        else:
            nb1 = nb2 = 0
            bp, ip = C.cast(None, U32P), C.cast(None, DP)
        self._check(lib().pofd2d_add_field(self.ptr, b, _dp(d1), _dp(d2),
                                           d1.size, bp, nb1, nb2, ip,
                                           sigma_base1, sigma_base2,
                                           like_norm, C.byref(i)),
                    "pofd2d_add_field")
        return i.value

    def prepare2d(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        self._check(lib().pofd2d_prepare(self.ptr, t.ctypes.data_as(FP),
                                         t.size), "pofd2d_prepare")

    def set_rrange2d(self, b, rng):
        r = _f64(rng)
        self._check(lib().pofd2d_set_rrange(self.ptr, b, _dp(r)),
                    "pofd2d_set_rrange")

    def beamset_info2d(self, b):
        out = np.zeros(14)
        self._check(lib().pofd2d_beamset_info(self.ptr, b, _dp(out)),
                    "pofd2d_beamset_info")
        keys = ["minR1", "maxR1", "minR2", "maxR2", "maxsig1", "maxsig2",
                "mind1", "maxd1", "mind2", "maxd2", "nfields", "like_offset0",
                "dflux1", "dflux2"]
        return dict(zip(keys, out))

    def loglike2d_batch(self, thetas):
        t = np.ascontiguousarray(np.atleast_2d(thetas), dtype=np.float32)
        w, stride = t.shape
        ll = np.zeros(w)
        st = np.zeros(w, dtype=np.int32)
        self._check(lib().pofd2d_loglike_batch(self.ptr, t.ctypes.data_as(FP),
                                               w, stride, _dp(ll),
                                               st.ctypes.data_as(I32P)),
                    "pofd2d_loglike_batch")
        return ll, st

    def loglike2d_batch_device(self, d_theta, nwalk, stride, d_ll, d_st,
                               stream=0):
        self._check(lib().pofd2d_loglike_batch_device(
            self.ptr, d_theta, nwalk, stride, d_ll, d_st, stream),
            "pofd2d_loglike_batch_device")

    def get_bonus2d(self, nwalk):
        """[nwalk, 5]: getNS, flux per area (band 1, 2), flux^2 per area
        (band 1, 2) per walker of the last batch."""
        out = np.zeros(5 * nwalk)
        self._check(lib().pofd2d_get_bonus(self.ptr, nwalk, _dp(out)),
                    "pofd2d_get_bonus")
        return out.reshape(nwalk, 5)

    def get_debug2d(self, walker, b=0, f=0):
        out = np.zeros(32)
        self._check(lib().pofd2d_get_debug(self.ptr, walker, b, f, _dp(out)),
                    "pofd2d_get_debug")
        return dict(zip(self.DEBUG_KEYS, out))

    def get_pd2d(self, walker, b=0, f=0, is_log=True):
        n = self.n
        pd = np.zeros(n * n)
        info = np.zeros(8)
        self._check(lib().pofd2d_get_pd(self.ptr, walker, b, f, int(is_log),
                                        _dp(pd), _dp(info)), "pofd2d_get_pd")
        keys = ["minflux1", "minflux2", "dflux1", "dflux2", "split1",
                "split2", "sgpos1", "sgpos2"]
        return pd.reshape(n, n), dict(zip(keys, info))



# ------------------------------------------------------------ field groups
class FieldGraph:
    """Several 1D contexts (fields of different nflux sharing theta) as one
    captured HIP graph whose launch writes the summed log likelihood."""

    def __init__(self, ctxs, d_theta, nwalk, stride, d_ll, d_st, stream):
        self.ctxs = list(ctxs)  # keep the contexts alive
        arr = (VP * len(self.ctxs))(*[c.ptr for c in self.ctxs])
        self.ptr = VP()
        rc = lib().pofd1d_graph_create(arr, len(self.ctxs), VP(d_theta),
                                       nwalk, stride, VP(d_ll), VP(d_st),
                                       VP(stream), C.byref(self.ptr))
        if rc:
            raise PofdError(rc, "pofd1d_graph_create failed")

    def launch(self, stream):
        rc = lib().pofd1d_graph_launch(self.ptr, VP(stream))
        if rc:
            raise PofdError(rc, "pofd1d_graph_launch failed")

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().pofd1d_graph_destroy(self.ptr)
            self.ptr = None


# ------------------------------------------------------------ sampler
TARGET_1D, TARGET_2D, TARGET_GAUSS = 1, 2, 3


def acor(x):
    """(tau, mean, sigma) of a series (affineChainSet::acor)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    m, sg, t = C.c_double(), C.c_double(), C.c_double()
    rc = lib().pofd_acor(x.ctypes.data_as(FP), x.size, C.byref(m),
                         C.byref(sg), C.byref(t))
    if rc:
        raise PofdError(rc, "pofd_acor: series too short")
    return t.value, m.value, sg.value


class Ensemble:
    """Device-resident affine stretch-move sampler over a likelihood
    context (or the Gaussian validation target)."""

    def __init__(self, ctx, target, nwalkers, npar, scalefac=2.0, seed=1):
        self.ptr = VP()
        self.ctx = ctx  # keep the context alive
        self.nwalk, self.npar = nwalkers, npar
        rc = lib().pofd_ens_create(C.byref(self.ptr),
                                   ctx.ptr if ctx is not None else None,
                                   target, nwalkers, npar, scalefac, seed)
        if rc:
            raise PofdError(rc, "pofd_ens_create failed")

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().pofd_ens_destroy(self.ptr)
            self.ptr = None

    def _check(self, rc, what):
        if rc:
            msg = lib().pofd_ens_last_error(self.ptr)
            raise PofdError(rc, "%s: %s" % (what, (msg or b"").decode()))

    def set_param_state(self, state):
        st = np.ascontiguousarray(state, dtype=np.uint8)
        self._check(lib().pofd_ens_set_param_state(
            self.ptr, st.ctypes.data_as(C.POINTER(C.c_uint8))),
            "pofd_ens_set_param_state")

    def set_exponent_mode(self, fit_dimension):
        """Acceptance exponent of z: the reference's nparams - nfixed - 1
        (fit_dimension False) or the fit parameters less one (True).
        Returns the exponent."""
        e = C.c_float()
        self._check(lib().pofd_ens_set_exponent_mode(self.ptr,
                                                     int(fit_dimension),
                                                     C.byref(e)),
                    "pofd_ens_set_exponent_mode")
        return e.value

    def set_limits(self, lower=None, upper=None):
        def arr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float32)
            self._keep = getattr(self, "_keep", []) + [a]
            return a.ctypes.data_as(FP)
        self._check(lib().pofd_ens_set_limits(self.ptr, arr(lower),
                                              arr(upper)),
                    "pofd_ens_set_limits")

    def set_gauss_target(self, mean, sigma):
        m, s = _f64(mean), _f64(sigma)
        self._check(lib().pofd_ens_set_gauss_target(self.ptr, _dp(m), _dp(s)),
                    "pofd_ens_set_gauss_target")

    def set_positions(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float32).reshape(
            self.nwalk, self.npar)
        self._check(lib().pofd_ens_set_positions(self.ptr, t.ctypes.data_as(FP)),
                    "pofd_ens_set_positions")

    def run(self, nsteps, temperature=1.0, thin=1):
        self._check(lib().pofd_ens_run(self.ptr, nsteps, temperature, thin),
                    "pofd_ens_run")

    def state(self):
        th = np.zeros((self.nwalk, self.npar), dtype=np.float32)
        ll = np.zeros(self.nwalk)
        na = np.zeros(self.nwalk, dtype=np.uint32)
        ne = np.zeros(self.nwalk, dtype=np.uint32)
        self._check(lib().pofd_ens_get_state(
            self.ptr, th.ctypes.data_as(FP), _dp(ll),
            na.ctypes.data_as(U32P), ne.ctypes.data_as(U32P)),
            "pofd_ens_get_state")
        return th, ll, na, ne

    def chain(self):
        n = C.c_uint32()
        self._check(lib().pofd_ens_chain_length(self.ptr, C.byref(n)),
                    "pofd_ens_chain_length")
        th = np.zeros((n.value, self.nwalk, self.npar), dtype=np.float32)
        ll = np.zeros((n.value, self.nwalk))
        if n.value:
            self._check(lib().pofd_ens_get_chain(
                self.ptr, 0, n.value, th.ctypes.data_as(FP), _dp(ll)),
                "pofd_ens_get_chain")
        return th, ll

    def clear_chain(self):
        self._check(lib().pofd_ens_clear_chain(self.ptr), "pofd_ens_clear_chain")

    def acor(self, param):
        t, m, s = C.c_double(), C.c_double(), C.c_double()
        self._check(lib().pofd_ens_acor(self.ptr, param, C.byref(t),
                                        C.byref(m), C.byref(s)),
                    "pofd_ens_acor")
        return t.value, m.value, s.value

    def sample_z(self, n):
        out = np.zeros(n)
        self._check(lib().pofd_ens_sample_z(self.ptr, n, _dp(out)),
                    "pofd_ens_sample_z")
        return out

    def set_surrogate(self, mlp, param_index=None, gate=0.5,
                      holdout_rms=None):
        """Delayed acceptance with the network mlp over the parameters
        param_index; switched on only when holdout_rms <= gate (the
        network's last training holdout error by default).  mlp None
        switches it off.  Returns whether it is on."""
        self._mlp = mlp  # keep the network alive
        if mlp is None:
            self._check(lib().pofd_ens_set_surrogate(self.ptr, None, None, 0,
                                                     gate, 0.0, None),
                        "pofd_ens_set_surrogate")
            return False
        idx = np.ascontiguousarray(param_index, dtype=np.uint32)
        if holdout_rms is None:
            holdout_rms = mlp.holdout_rms
        on = C.c_int()
        self._check(lib().pofd_ens_set_surrogate(
            self.ptr, mlp.ptr, idx.ctypes.data_as(U32P), idx.size, gate,
            holdout_rms, C.byref(on)), "pofd_ens_set_surrogate")
        return bool(on.value)

    def set_surrogate_steps(self, nsteps):
        """Surrogate stretch moves per walker and half-step while the
        surrogate is on: 1 is delayed acceptance; more runs an inner chain
        on the surrogate with one exact correction at its end (the
        surrogate transition).  1 to 64."""
        self._check(lib().pofd_ens_set_surrogate_steps(self.ptr, int(nsteps)),
                    "pofd_ens_set_surrogate_steps")

    def counts(self):
        """(proposals, exact likelihood evaluations) since set_positions."""
        a, b = C.c_uint64(), C.c_uint64()
        self._check(lib().pofd_ens_counts(self.ptr, C.byref(a), C.byref(b)),
                    "pofd_ens_counts")
        return a.value, b.value


# ------------------------------------------------------------ surrogate
class Mlp:
    """Log-likelihood surrogate nin -> H -> H -> H -> 1 (GELU, bfloat16
    weights and activations, float accumulation), trained on the GPU."""

    def __init__(self, nin, hidden=256, seed=1):
        self.ptr = VP()
        self.nin, self.hidden = nin, hidden
        self.train_rms = self.holdout_rms = float("nan")
        rc = lib().pofd_mlp_create(C.byref(self.ptr), nin, hidden, seed)
        if rc:
            raise PofdError(rc, "pofd_mlp_create failed")
        n = C.c_uint32()
        self._check(lib().pofd_mlp_nparams(self.ptr, C.byref(n)),
                    "pofd_mlp_nparams")
        self.nparams = n.value
        off = (C.c_uint64 * 8)()
        self._check(lib().pofd_mlp_layout(self.ptr, off), "pofd_mlp_layout")
        self.offsets = [int(v) for v in off]

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().pofd_mlp_destroy(self.ptr)
            self.ptr = None

    def _check(self, rc, what):
        if rc:
            msg = lib().pofd_mlp_last_error(self.ptr)
            raise PofdError(rc, "%s: %s" % (what, (msg or b"").decode()))

    def params(self):
        w = np.zeros(self.nparams, dtype=np.float32)
        self._check(lib().pofd_mlp_get_params(self.ptr, w.ctypes.data_as(FP)),
                    "pofd_mlp_get_params")
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.nparams
        self._check(lib().pofd_mlp_set_params(self.ptr, w.ctypes.data_as(FP)),
                    "pofd_mlp_set_params")

    def set_standardization(self, xmean, xstd, ymean, ystd):
        m = np.ascontiguousarray(xmean, dtype=np.float32)
        s = np.ascontiguousarray(xstd, dtype=np.float32)
        self._check(lib().pofd_mlp_set_standardization(
            self.ptr, m.ctypes.data_as(FP), s.ctypes.data_as(FP),
            float(ymean), float(ystd)), "pofd_mlp_set_standardization")

    def set_whitening(self, xmean, wmat, ymean, ystd):
        m = np.ascontiguousarray(xmean, dtype=np.float32)
        w = np.ascontiguousarray(wmat, dtype=np.float32).reshape(self.nin,
                                                                 self.nin)
        self._check(lib().pofd_mlp_set_whitening(
            self.ptr, m.ctypes.data_as(FP), w.ctypes.data_as(FP),
            float(ymean), float(ystd)), "pofd_mlp_set_whitening")

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.nin)
        out = np.zeros(x.shape[0])
        self._check(lib().pofd_mlp_predict(self.ptr, x.ctypes.data_as(FP),
                                           x.shape[0], _dp(out)),
                    "pofd_mlp_predict")
        return out

    def loss_grad(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.nin)
        y = _f64(y)
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = C.c_double()
        self._check(lib().pofd_mlp_loss_grad(
            self.ptr, x.ctypes.data_as(FP), _dp(y), x.shape[0],
            g.ctypes.data_as(FP), C.byref(loss)), "pofd_mlp_loss_grad")
        return loss.value, g

    def train(self, x, y, epochs=100, batch=256, lr=1e-3, holdout_frac=0.1,
              seed=1):
        """Adam training; returns (train RMS, holdout RMS) in the units of
        y."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.nin)
        y = _f64(y)
        tr, ho = C.c_double(), C.c_double()
        self._check(lib().pofd_mlp_train(
            self.ptr, x.ctypes.data_as(FP), _dp(y), x.shape[0], epochs, batch,
            lr, holdout_frac, seed, C.byref(tr), C.byref(ho)),
            "pofd_mlp_train")
        self.train_rms, self.holdout_rms = tr.value, ho.value
        return tr.value, ho.value
