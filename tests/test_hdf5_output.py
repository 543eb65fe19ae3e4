"""HDF5 output in the reference's layout (include/pofd_affine_amd_hdf5.hpp),
written through the C++ headers and read back through the HDF5 C API by
tests/cpp/hdf5_check: PD::writeToHDF5 (src/PD.cc:489-526),
PDDouble::writeToHDF5 (src/PDDouble.cc:912-966), the chain layout of
affineChainSet::writeToHDF5Handle (src/affineChainSet.cc:1306-1450),
initFileKnots::writeToHDF5Handle (src/initFileKnots.cc:627-655) on the
reference's initfile_test1D.txt, and specFile::writeToHDF5Handle
(src/specFile.cc:406-416).  No GPU: these writers are host code."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import REFDATA

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "hdf5_check")


def h5dump(path):
    """{'groups': set, 'attrs': {path@name: (kind, values)}, 'data':
    {path: (kind, dims, values)}} of an HDF5 file."""
    out = subprocess.run([EXE, "dump", str(path)], capture_output=True,
                         text=True, timeout=60, check=True).stdout
    res = {"groups": set(), "attrs": {}, "data": {}}

    def conv(kind, toks):
        if kind[0] == "s":
            return [t[1:-1] for t in toks]
        if kind[0] == "f":
            return np.array([float(t) for t in toks])
        return np.array([int(t) for t in toks])

    for line in out.splitlines():
        w = line.split(" ")
        if w[0] == "G":
            res["groups"].add(w[1])
        elif w[0] == "A":
            res["attrs"][w[1]] = (w[2], conv(w[2], w[4:]))
        elif w[0] == "D":
            rank = int(w[3])
            dims = tuple(int(t) for t in w[4:4 + rank])
            res["data"][w[1]] = (w[2], dims, conv(w[2], w[5 + rank:]))
    return res


def _att(d, key):
    return d["attrs"][key][1]


@pytest.fixture(scope="module")
def written(tmp_path_factory):
    assert os.path.exists(EXE), "build tests/cpp first (__graft_entry__.build)"
    d = tmp_path_factory.mktemp("h5")
    out = subprocess.run([EXE, "write", str(d),
                          os.path.join(REFDATA, "initfile_test1D.txt")],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    return d, out.stdout


def test_pd_layout(written):
    d, _ = written
    h = h5dump(d / "pd.h5")
    assert h["attrs"]["/@IsLog"] == ("u1", [1])
    assert h["attrs"]["/@dFlux"][0] == "f8" and _att(h, "/@dFlux")[0] == 0.125
    assert _att(h, "/@MinFlux")[0] == -0.25
    kind, dims, flux = h["data"]["/Flux"]
    assert kind == "f8" and dims == (5,)
    np.testing.assert_array_equal(flux, -0.25 + 0.125 * np.arange(5))
    np.testing.assert_array_equal(h["data"]["/PD"][2],
                                  [0.5, 1.5, -2.0, 3.25, 4.0])


def test_pddouble_layout(written):
    d, _ = written
    h = h5dump(d / "pd2.h5")
    assert _att(h, "/@IsLog")[0] == 0
    assert _att(h, "/@dFlux1")[0] == 0.5 and _att(h, "/@dFlux2")[0] == 0.25
    assert _att(h, "/@MinFlux1")[0] == -1.0 and _att(h, "/@MinFlux2")[0] == 2.0
    np.testing.assert_array_equal(h["data"]["/Flux1"][2], [-1.0, -0.5, 0.0])
    np.testing.assert_array_equal(h["data"]["/Flux2"][2],
                                  [2.0, 2.25, 2.5, 2.75])
    kind, dims, pd = h["data"]["/PD"]
    assert kind == "f8" and dims == (3, 4)
    np.testing.assert_array_equal(pd, 0.1 * np.arange(12))


def test_chain_layout(written):
    """Chain [nwalkers][nsteps][nparams] float, Likelihood [nwalkers][nsteps]
    double, from a sampler-order chain theta = 100 s + 10 w + p,
    logL = -(10 s + w)."""
    d, _ = written
    h = h5dump(d / "chain.h5")
    kind, dims, ch = h["data"]["/Chain"]
    assert kind == "f4" and dims == (4, 3, 2)
    w, s, p = np.meshgrid(np.arange(4), np.arange(3), np.arange(2),
                          indexing="ij")
    np.testing.assert_array_equal(ch.reshape(dims), 100 * s + 10 * w + p)
    for k, v in (("NWalkers", 4), ("NSteps", 3), ("NParams", 2)):
        assert h["attrs"]["/Chain@" + k] == ("u4", [v])
    kind, dims, ll = h["data"]["/Likelihood"]
    assert kind == "f8" and dims == (4, 3)
    w, s = np.meshgrid(np.arange(4), np.arange(3), indexing="ij")
    np.testing.assert_array_equal(ll.reshape(dims), -(10 * s + w))
    assert _att(h, "/Likelihood@NWalkers")[0] == 4
    assert "/Likelihood@NParams" not in h["attrs"]


def test_initfile_and_spec_layout(written):
    """The reference's initfile_test1D.txt (ranges and limits) as
    initFileKnots::writeToHDF5Handle writes it; limits a knot lacks are
    NaN."""
    d, _ = written
    h = h5dump(d / "meta.h5")
    assert {"/ParamInfo", "/LikelihoodParams"} <= h["groups"]
    for k in ("HasInitParamRange", "HasParamLowerLimits",
              "HasParamUpperLimits"):
        assert h["attrs"]["/ParamInfo@" + k] == ("u1", [1])
    rows = []
    with open(os.path.join(REFDATA, "initfile_test1D.txt")) as f:
        for line in f:
            t = line.split()
            if len(t) >= 3 and not line.lstrip().startswith("#"):
                rows.append([float(x) for x in t])
    n = len(rows)
    rng = h["data"]["/ParamInfo/ParamInitRange"][2]
    np.testing.assert_array_equal(rng, [r[2] for r in rows])
    has_lo = h["data"]["/ParamInfo/ParamHasLowerlim"]
    assert has_lo[0] == "u1" and has_lo[1] == (n,)
    lo = h["data"]["/ParamInfo/ParamLowerlim"][2]
    hi = h["data"]["/ParamInfo/ParamUpperlim"][2]
    has_hi = h["data"]["/ParamInfo/ParamHasUpperlim"][2]
    for i, r in enumerate(rows):
        fixed = r[2] == 0
        want_lo = len(r) > 3 and not fixed
        assert bool(has_lo[2][i]) == want_lo, i
        if want_lo:
            assert lo[i] == r[3]
        else:
            assert math.isnan(lo[i])
        want_hi = len(r) > 4 and not fixed
        assert bool(has_hi[i]) == want_hi, i
        if want_hi:
            assert hi[i] == r[4]
    lp = "/LikelihoodParams/"
    assert h["data"][lp + "DataFiles"][2] == ["map_a.fits", "map_b.fits"]
    assert h["data"][lp + "BeamFiles"][2] == ["beam.fits", "beam.fits"]
    np.testing.assert_array_equal(h["data"][lp + "InstrumentSigma"][2],
                                  [0.002, 0.003])
    np.testing.assert_array_equal(
        h["data"][lp + "LikelihoodNormalization"][2], [1.0, 0.5])
    assert h["attrs"]["/LikelihoodParams@Names"] == \
        ("s0", ["alpha", "", "gamma"])


def test_unwritable_path_raises(written):
    _, stdout = written
    assert "unwritable PD writeToHDF5" in stdout
