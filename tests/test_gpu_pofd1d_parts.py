"""The 1D transforms split over 1, 2 or 4 workgroups per walker
(pofd1d_set_parts).  Every sum the split touches (the kept rows and
moments of the forward transform, the P(D) sums of the inverse) runs over
four fixed chunk groups, each summed in chunk order and combined as
(G0 + G1) + (G2 + G3), and a part holds whole groups: a walker's log
likelihood and P(D) are the same bits for every part count, and so for
every batch size.  The full-size values are also held to the exact-math
value of the reference's algorithm (the extended-precision oracle) within
the 1e-6 parity bar."""
import os

import numpy as np
import pytest

from conftest import REFDATA
from oracle_side import build_oracle
from pofd_affine_amd import fitsio, hiprt, synth, workload

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-parts-case] " + msg, flush=True)


def _by_parts(ctx, th):
    out = {}
    for parts in (1, 2, 4):
        ctx.set_parts(parts)
        ll, st = ctx.loglike_batch(th)
        assert np.all(st == 0), (parts, st)
        # unbinned fields keep only log2 of the P(D)
        pd = 2.0 ** ctx.get_pd(len(th) - 1, is_log=True)[0]
        out[parts] = (ll, pd)
    ctx.set_parts(0)
    return out


@pytest.mark.parametrize("binned", [True, False])
def test_parts_full_size(binned):
    _say("full size by parts, binned=%s" % binned)
    # sources inside the knot range (<= 450 mJy), as test_config1d_in_support
    inp = workload.make_inputs(npix_side=1000, seed=42, smax=0.45)
    th = synth.walker_thetas(24, seed=99)
    ctx = workload.build_context(inp, 131072, 2048, binned=binned)
    out = _by_parts(ctx, th)
    for parts in (2, 4):
        np.testing.assert_array_equal(out[parts][0], out[1][0])
        np.testing.assert_array_equal(out[parts][1], out[1][1])
    L = build_oracle(inp, 131072, 2048, binned, th[0])
    L.set_precise(1)
    ext, rst = L.loglike_batch(th, 16)
    assert np.all(rst == 0)
    np.testing.assert_allclose(out[1][0], ext, rtol=0, atol=1e-6)


def test_parts_negative_beam():
    _say("negative beam by parts")
    pix, ps = fitsio.read_beam(os.path.join(REFDATA,
                                            "band12_beam_f100.fits[0]"))
    inp = dict(workload.make_inputs(npix_side=200, seed=7), beam=pix,
               pixsize=ps)
    th = synth.walker_thetas(6, seed=13)
    ctx = workload.build_context(inp, 32768, 512, binned=True)
    out = _by_parts(ctx, th)
    L = build_oracle(inp, 32768, 512, True, th[0])
    ref, rst = L.loglike_batch(th, 6)
    assert np.all(rst == 0)
    for parts in (2, 4):
        np.testing.assert_array_equal(out[parts][0], out[1][0])
        np.testing.assert_array_equal(out[parts][1], out[1][1])
    np.testing.assert_allclose(out[1][0], ref, rtol=0, atol=1e-6)


def test_parts_auto_choice_and_limits():
    """The automatic choice splits small batches; a forced count that does
    not fit the workspace's slots is reduced, and a bad count (3, 5)
    refused.  Every split gives the same bits."""
    _say("parts auto choice")
    inp = workload.make_inputs(npix_side=200, seed=7)
    ctx = workload.build_context(inp, 32768, 512, binned=True)
    th = synth.walker_thetas(200, seed=5)
    ll_big, st = ctx.loglike_batch(th)               # 1 part per walker
    assert np.all(st == 0)
    ll_small, st = ctx.loglike_batch(th[:40])        # 4 parts per walker
    assert np.all(st == 0)
    np.testing.assert_array_equal(ll_small, ll_big[:40])
    ll_mid, st = ctx.loglike_batch(th[:70])          # 256 / 70 = 3 -> 2
    assert np.all(st == 0)
    np.testing.assert_array_equal(ll_mid, ll_big[:70])
    ctx.set_parts(4)
    ll_forced, st = ctx.loglike_batch(th)            # reduced to fit
    assert np.all(st == 0)
    np.testing.assert_array_equal(ll_forced, ll_big)
    ctx.set_parts(0)
    for bad in (3, 5):
        with pytest.raises(Exception):
            ctx.set_parts(bad)


@pytest.mark.parametrize("count", [0, 1, 37, 70, 150])
def test_device_count_batch(count):
    """pofd1d_loglike_batch_device_count: the batch size read from device
    memory when the kernels run; rows past it are left untouched."""
    _say("device count %d" % count)
    inp = workload.make_inputs(npix_side=200, seed=7)
    ctx = workload.build_context(inp, 32768, 512, binned=True)
    nmax = 160
    th = synth.walker_thetas(nmax, seed=17)
    ref, st = ctx.loglike_batch(th[:max(count, 1)])
    assert np.all(st == 0)
    d_th = hiprt.DeviceBuffer(th.nbytes)
    d_th.upload(th)
    d_n = hiprt.DeviceBuffer(4)
    d_n.upload(np.array([count], dtype=np.int32))
    d_ll = hiprt.DeviceBuffer(nmax * 8)
    d_st = hiprt.DeviceBuffer(nmax * 4)
    d_ll.upload(np.full(nmax, 123.0))
    d_st.upload(np.full(nmax, -7, dtype=np.int32))
    ctx.loglike_batch_device_count(d_th.value, nmax, d_n.value, th.shape[1],
                                   d_ll.value, d_st.value)
    hiprt.device_synchronize()
    ll = np.zeros(nmax)
    stv = np.zeros(nmax, dtype=np.int32)
    d_ll.download(ll)
    d_st.download(stv)
    assert np.all(stv[:count] == 0)
    # the same part count as a host batch of this size
    np.testing.assert_array_equal(ll[:count], ref[:count])
    assert np.all(ll[count:] == 123.0) and np.all(stv[count:] == -7)


def test_autotune_parts():
    """pofd1d_autotune_parts: every measured choice is an allowed split, a
    batch then runs exactly as with that split forced (unmeasured sizes as
    the nearest measured one, device-count batches as host ones), and
    tuning with no sizes restores the size rule."""
    _say("autotune parts")
    inp = workload.make_inputs(npix_side=200, seed=7)
    ctx = workload.build_context(inp, 32768, 512, binned=True)
    th = synth.walker_thetas(200, seed=23)
    rule = {}
    for n in (8, 37, 40, 50, 70, 130, 200):
        rule[n], st = ctx.loglike_batch(th[:n])
        assert np.all(st == 0)
    sizes = [8, 40, 70, 130, 200]
    chosen = ctx.autotune_parts(th, sizes, reps=3)
    _say("chosen %s" % list(chosen))
    choice = {}
    for n, p in zip(sizes, chosen):
        assert p in (1, 2, 4) and p <= min(4, 256 // n), (n, p)
        choice[n] = int(p)
    choice[50] = choice[40]     # nearest measured size
    choice[37] = choice[40]
    for n, p in choice.items():
        tuned, st = ctx.loglike_batch(th[:n])
        assert np.all(st == 0)
        ctx.set_parts(p)
        forced, st = ctx.loglike_batch(th[:n])
        ctx.set_parts(0)
        np.testing.assert_array_equal(tuned, forced)
        np.testing.assert_array_equal(tuned, rule[n])
    # a device-count batch takes the same measured split as a host batch
    nmax, count = 160, 37
    d_th = hiprt.DeviceBuffer(th[:nmax].nbytes)
    d_th.upload(np.ascontiguousarray(th[:nmax]))
    d_n = hiprt.DeviceBuffer(4)
    d_n.upload(np.array([count], dtype=np.int32))
    d_ll = hiprt.DeviceBuffer(nmax * 8)
    d_st = hiprt.DeviceBuffer(nmax * 4)
    ctx.loglike_batch_device_count(d_th.value, nmax, d_n.value, th.shape[1],
                                   d_ll.value, d_st.value)
    hiprt.device_synchronize()
    ll = np.zeros(nmax)
    d_ll.download(ll)
    ref, st = ctx.loglike_batch(th[:count])
    np.testing.assert_array_equal(ll[:count], ref)
    # no sizes: the size rule again (40 walkers -> 4 parts)
    ctx.autotune_parts(th, [], reps=1)
    again, st = ctx.loglike_batch(th[:40])
    ctx.set_parts(4)
    forced, st2 = ctx.loglike_batch(th[:40])
    ctx.set_parts(0)
    np.testing.assert_array_equal(again, forced)
