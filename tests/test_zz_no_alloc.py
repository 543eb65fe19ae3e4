"""The allocation-free batch entry point (pofd1d_loglike_batch_device_noalloc)
against pofd1d_loglike_batch_device through the C ABI: equal results, bit
for bit, on two calls in a row (the second with fewer walkers), with two
beam sets in the context and with a split batch (several workgroups per
walker); a batch beyond the device scratch is refused with POFD_E_ARG."""
import os

import numpy as np
import pytest

from conftest import REFDATA
from pofd_affine_amd import _lib, fitsio, hiprt, synth, workload

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-noalloc-case] " + msg, flush=True)


def _run(ctx, th, noalloc):
    W, stride = th.shape
    d_th = hiprt.DeviceBuffer(th.nbytes)
    d_ll = hiprt.DeviceBuffer(W * 8)
    d_st = hiprt.DeviceBuffer(W * 4)
    d_th.upload(np.ascontiguousarray(th))
    stream = hiprt.Stream()
    f = ctx.loglike_batch_device_noalloc if noalloc else ctx.loglike_batch_device
    f(d_th.value, W, stride, d_ll.value, d_st.value, stream.ptr.value)
    stream.synchronize()
    ll = np.zeros(W)
    st = np.zeros(W, dtype=np.int32)
    d_ll.download(ll)
    d_st.download(st)
    return ll, st


def test_noalloc_matches_allocating_entry():
    _say("nflux 2^17, 48 then 8 walkers, allocating vs allocation-free")
    inp = workload.make_inputs(npix_side=300, seed=21)
    ctx = workload.build_context(inp, 131072, 2048, binned=True)
    for W, seed in ((48, 3), (8, 4)):
        th = synth.walker_thetas(W, seed=seed)
        a, sa = _run(ctx, th, noalloc=True)
        b, sb = _run(ctx, th, noalloc=False)
        assert np.all(sa == 0) and np.array_equal(sa, sb)
        assert np.array_equal(a, b), np.max(np.abs(a - b))
        _say("%d walkers equal" % W)


def test_noalloc_two_beamsets_unbinned():
    _say("two beam sets (one two-signed), unbinned and binned, 2^15")
    gb, gps = synth.gaussian_beam()
    nb, nps = fitsio.read_beam(os.path.join(REFDATA,
                                            "band12_beam_f100.fits[0]"))
    ctx = _lib.Context(0)
    ctx.configure(32768, 512, synth.knot_positions(), 0.0)
    beams = []
    for k, (beam, ps) in enumerate(((gb.ravel(), gps), (nb, nps))):
        bm = _lib.Beam(beam, ps, 1e-6, 120)
        beams.append(bm)
        b = ctx.add_beamset(bm)
        d, _ = synth.make_map(150, 0.003, seed=30 + k)
        if k == 0:
            ctx.add_field(b, pix=d, sigma_base=0.003)
        else:
            bins, c0, dl = _lib.data_bin(d, 500)
            ctx.add_field(b, bins=bins, bincent0=c0, bindelta=dl,
                          sigma_base=0.003, npix=d.size)
    ctx.prepare()
    ctx.set_priors(sigma_width=0.1)
    for W, seed in ((20, 5), (6, 6)):
        th = synth.walker_thetas(W, seed=seed)
        a, sa = _run(ctx, th, noalloc=True)
        b, sb = _run(ctx, th, noalloc=False)
        assert np.array_equal(sa, sb) and np.all(sa == 0)
        assert np.array_equal(a, b)
    ctx.set_parts(2)
    th = synth.walker_thetas(16, seed=8)
    a, _ = _run(ctx, th, noalloc=True)
    b, _ = _run(ctx, th, noalloc=False)
    assert np.array_equal(a, b)


def test_noalloc_refuses_beyond_scratch():
    _say("a batch beyond the scratch is refused")
    inp = workload.make_inputs(npix_side=200, seed=9)
    ctx = workload.build_context(inp, 131072, 2048, binned=True)
    th = synth.walker_thetas(256, seed=3)
    with pytest.raises(_lib.PofdError) as ei:
        _run(ctx, th, noalloc=True)
    assert ei.value.code == 5  # POFD_E_ARG
    assert "MiB device scratch" in str(ei.value)
