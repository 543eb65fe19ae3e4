"""The reference's input and output formats through
include/pofd_affine_amd_io.hpp (tests/cpp/io_check.cpp), against the known
answers of the reference's own tests on its own fixtures:
test/specFile_test.cc (specfile_test1D.txt, specfile_test2D.txt) and
test/initFile_test.cc (fiducial_1Dmodel.txt, initfile_test1D.txt,
initfile_test2D.txt); the NR3 generator against its published first
outputs' properties; and the text chain layout of
affineChainSet::writeToFile (src/affineChainSet.cc:406-423)."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import REFDATA

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
EXE = os.path.join(CPP, "_build", "io_check")


@pytest.fixture(scope="module")
def io(tmp_path_factory):
    # the check program is header-only code over the library; (re)build it
    # when missing or stale
    subprocess.run(["make", "-C", CPP, "_build/io_check"], check=True,
                   capture_output=True, timeout=300)
    chain = str(tmp_path_factory.mktemp("io") / "chain.txt")
    out = subprocess.run([EXE, REFDATA, chain], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    vals = {}
    blocks = {}
    for line in out.stdout.splitlines():
        w = line.split()
        if w[0].endswith(".k"):
            blocks.setdefault(w[0], []).append(dict(zip(w[2::2], w[3::2])))
        else:
            vals[w[0]] = w[1:]
    return vals, blocks, chain


def _f(v):
    return [float(x) for x in v]


def test_spec_file_1d(io):
    """test/specFile_test.cc:12-109"""
    v, _, _ = io
    assert v["s1.datafiles"] == ["file1.fits", "file2.fits"]
    assert _f(v["s1.sigmas"]) == pytest.approx([0.0014, 0.002], abs=1e-4)
    assert v["s1.psffiles"] == ["beam1.fits", "beam3.fits"]
    assert _f(v["s1.like_norm"]) == pytest.approx([1.0, 1.1], abs=1e-4)
    assert v["s1.bin_data"] == ["1"] and v["s1.nbins"] == ["500"]
    assert v["s1.mean_sub"] == ["1"] and v["s1.ignore_mask"] == ["0"]
    assert v["s1.fftsize"] == ["2048"] and v["s1.ninterp"] == ["1200"]
    assert v["s1.beam_histogram"] == ["1"]
    # fit_sigma = no, then sigmaprior= turns it on
    assert v["s1.fit_sigma"] == ["1"] and v["s1.has_sigprior"] == ["1"]
    assert float(v["s1.sigprior_stdev"][0]) == pytest.approx(0.15, abs=1e-4)
    assert float(v["s1.exp_conf"][0]) == 0.0
    assert float(v["s1.minbeamval"][0]) == pytest.approx(1e-6, rel=1e-6)
    assert v["s1.nbeamhist"] == ["120"]
    assert v["s1.has_cfirbprior"] == ["1"]
    assert float(v["s1.cfirbprior_mean"][0]) == pytest.approx(1.05, abs=1e-4)
    assert float(v["s1.cfirbprior_stdev"][0]) == pytest.approx(0.1, abs=1e-4)
    assert v["s1.has_poissonprior"] == ["0"]
    assert float(v["s1.regularization_alpha"][0]) == 0.0
    assert v["s1.has_wisdom_file"] == ["1"]
    assert v["s1.wisdom_file"] == ["a_wisdom_file.txt"]
    assert v["s1.verbosity"] == ["1"]   # verbose= true, ultraverbose= false


def test_spec_file_2d(io):
    """test/specFile_test.cc:112-240"""
    v, _, _ = io
    assert v["s2.datafiles1"] == ["file1_1.fits", "file2.fits"]
    assert v["s2.datafiles2"] == ["file1_2.fits", "file4.fits"]
    assert _f(v["s2.sigmas1"]) == pytest.approx([0.0014, 0.002], abs=1e-6)
    assert _f(v["s2.sigmas2"]) == pytest.approx([0.003, 0.04], abs=1e-6)
    assert v["s2.psffiles1"] == ["beam1.fits", "beam3.fits"]
    assert v["s2.psffiles2"] == ["beam5.fits", "funny_beam.fits"]
    assert _f(v["s2.like_norm"]) == pytest.approx([1.0, 1.1], abs=1e-6)
    assert v["s2.bin_data"] == ["1"] and v["s2.nbins"] == ["300"]
    assert v["s2.mean_sub"] == ["1"] and v["s2.ignore_mask"] == ["1"]
    assert v["s2.fftsize"] == ["2048"] and v["s2.nedge"] == ["100"]
    assert v["s2.beam_histogram"] == ["1"]
    assert v["s2.fit_sigma1"] == ["1"] and v["s2.fit_sigma2"] == ["0"]
    assert v["s2.has_sigprior1"] == ["1"] and v["s2.has_sigprior2"] == ["0"]
    assert float(v["s2.sigprior_stdev1"][0]) == pytest.approx(0.15)
    assert float(v["s2.exp_conf1"][0]) == 0.0
    assert float(v["s2.exp_conf2"][0]) == pytest.approx(0.006)
    assert v["s2.has_cfirbprior1"] == ["0"] and v["s2.has_cfirbprior2"] == ["1"]
    assert float(v["s2.cfirbprior_mean2"][0]) == pytest.approx(1.05)
    assert float(v["s2.cfirbprior_stdev2"][0]) == pytest.approx(0.1)
    assert v["s2.has_poissonprior1"] == ["1"]
    assert v["s2.has_poissonprior2"] == ["0"]
    assert float(v["s2.poissonprior_mean1"][0]) == pytest.approx(0.5)
    assert float(v["s2.poissonprior_stdev1"][0]) == pytest.approx(0.2)
    assert v["s2.wisdom_file"] == ["a_wisdom_file.txt"]
    assert v["s2.verbosity"] == ["1"]


def test_malformed_spec_lines_refused(io):
    v, _, _ = io
    assert v["bad.refused"] == ["4"]


def test_init_file_fiducial(io):
    """test/initFile_test.cc:14-59: positions, values, nothing fixed."""
    v, b, _ = io
    assert v["i0.nknots"] == ["10"]
    assert _f(v["i0.pos"]) == pytest.approx(
        [0.0001, 0.002, 0.005, 0.010, 0.02, 0.045, 0.100, 0.2, 0.45, 1.0])
    assert _f(v["i0.val"]) == pytest.approx(
        [10.75, 7.0, 6.3, 5.9, 5.1, 4.0, 2.6, 1.4, 0.6, -0.4])
    assert all(k["fixed"] == "0" for k in b["i0.k"])


KNOTPOS = [0.002, 0.005, 0.045, 0.100, 0.45, 1.0]
KNOTVAL = [7.0, 6.3, 4.0, 2.6, 0.6, -0.4]
FIXED = [0, 0, 0, 0, 0, 1]
RANGE = [0.02, 0.03, 0.02, 0.01, 0.05, 0.0]
HASLOW = [1, 0, 1, 0, 1, 0]
LOW = [3.0, 0, 1.0, 0, -2.0, 0]
HASUP = [1, 0, 0, 0, 1, 0]
UP = [8.0, 0.0, 0.0, 0.0, 3.0, 0.0]


def _check_block(rows, fixed, rng, haslow, low, hasup, up):
    for i, k in enumerate(rows):
        assert int(k["fixed"]) == fixed[i], i
        assert float(k["range"]) == pytest.approx(rng[i]), i
        assert int(k["haslow"]) == haslow[i], i
        if haslow[i]:
            assert float(k["low"]) == pytest.approx(low[i]), i
        assert int(k["hasup"]) == hasup[i], i
        if hasup[i]:
            assert float(k["up"]) == pytest.approx(up[i]), i


def test_init_file_limits(io):
    """test/initFile_test.cc:61-129: ranges, fixed knots and limits."""
    v, b, _ = io
    assert v["i1.nknots"] == ["6"]
    assert _f(v["i1.pos"]) == pytest.approx(KNOTPOS)
    assert _f(v["i1.val"]) == pytest.approx(KNOTVAL)
    _check_block(b["i1.k"], FIXED, RANGE, HASLOW, LOW, HASUP, UP)
    # generateRandomKnotValues: 20000 draws, all valid, inside the ranges,
    # the first knot spanning its range [6.99, 7.01)
    assert v["i1.draw_bad"] == ["0"]
    assert float(v["i1.draw0_min"][0]) == pytest.approx(6.99, abs=2e-5)
    assert float(v["i1.draw0_max"][0]) == pytest.approx(7.01, abs=2e-5)
    assert v["i1.valid_low"] == ["0"] and v["i1.valid_ok"] == ["1"]


def test_init_file_2d(io):
    """test/initFile_test.cc:131-235."""
    v, b, _ = io
    assert (v["i2.nknots"], v["i2.nsigmas"], v["i2.noffsets"],
            v["i2.ntot"]) == (["6"], ["3"], ["2"], ["11"])
    assert _f(v["i2.pos"]) == pytest.approx(KNOTPOS)
    assert _f(v["i2.val"]) == pytest.approx(KNOTVAL)
    assert _f(v["i2.sigpos"]) == pytest.approx([0.004, 0.03, 0.10])
    assert _f(v["i2.sigval"]) == pytest.approx([0.3, 0.35, 0.45])
    assert _f(v["i2.offpos"]) == pytest.approx([0.004, 0.5])
    assert _f(v["i2.offval"]) == pytest.approx([0.2, 0.3])
    rows = b["i2.k"]
    _check_block(rows[:6], FIXED, RANGE, HASLOW, LOW, HASUP, UP)
    _check_block(rows[6:9], [0, 0, 1], [0.03, 0.03, 0.0], [0, 1, 0],
                 [0.0, 0.1, 0.0], [0, 1, 0], [0.0, 0.6, 0.0])
    # offsets: the first fixed (range 0), the second with a lower limit
    _check_block(rows[9:], [1, 0], [0.0, 0.02], [0, 1], [0.0, 0.01],
                 [0, 0], [0.0, 0.0])


def _nr3_ran(seed, n):
    """The generator of Numerical Recipes 3rd ed. section 7.1.3, written
    out in Python integers (the published algorithm) to check the C++
    one."""
    M = (1 << 64) - 1
    st = {"u": 0, "v": 4101842887655102017, "w": 1}

    def int64():
        st["u"] = (st["u"] * 2862933555777941757 + 7046029254386353087) & M
        v = st["v"]
        v ^= v >> 17
        v = (v ^ (v << 31)) & M
        v ^= v >> 8
        st["v"] = v
        w = st["w"]
        st["w"] = (4294957665 * (w & 0xFFFFFFFF) + (w >> 32)) & M
        x = st["u"] ^ ((st["u"] << 21) & M)
        x ^= x >> 35
        x = (x ^ (x << 4)) & M
        return ((x + st["v"]) & M) ^ st["w"]

    st["u"] = seed ^ st["v"]
    int64()
    st["v"] = st["u"]
    int64()
    st["w"] = st["v"]
    int64()
    return [int64() for _ in range(n)]


def test_nr3_generator(io):
    v, _, _ = io
    assert [int(x) for x in v["ran.first"]] == _nr3_ran(1, 3)
    assert float(v["ran.mean"][0]) == pytest.approx(0.5, abs=1e-3)
    assert float(v["ran.var"][0]) == pytest.approx(1.0 / 12, abs=1e-3)


def test_text_chain_layout(io):
    """affineStepChunk::writeToStream: walker-major, one stored step per
    line, parameters in %e with the first in a 13-wide left-justified
    field, the log likelihood 15 wide with 8 digits, then the walker."""
    _, _, chain = io
    lines = open(chain).read().splitlines()
    assert len(lines) == 6
    assert lines[0] == "-1.000000e+00 -7.500000e-01 -5.000000e-01 " \
                       "-1.00000000e+03 0"
    assert lines[1] == "5.000000e-01  7.500000e-01 1.000000e+00 " \
                       "-1.00100000e+03 0"
    a = np.loadtxt(chain)
    # theta[s][w][k] = 0.25 (6 s + 3 w + k) - 1, logL[s][w] = -1000 - (2s + w)/2
    for row, (w, s) in zip(a, [(w, s) for w in range(2) for s in range(3)]):
        want = [0.25 * (6 * s + 3 * w + k) - 1.0 for k in range(3)]
        assert list(row[:3]) == pytest.approx(want)
        assert row[3] == pytest.approx(-1000.0 - 0.5 * (2 * s + w))
        assert int(row[4]) == w
    assert not math.isnan(a.sum())
