"""Fields of different nflux sharing theta as one captured HIP graph
(pofd1d_graph_create / pofd1d_graph_launch): every replay equals the sum of
the fields' eager batch log likelihoods, with eager work and new theta
between the replays.  The graph adds a field's prior after the running sum
rather than to that field's own value, so the two sums may differ in the
last bit (1e-9 absolute here, relative 1e-14)."""
import numpy as np
import pytest

from pofd_affine_amd import _lib, hiprt, synth, workload

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-graph-case] " + msg, flush=True)


def test_field_graph_replays_match_eager():
    _say("three fields, nflux 2^15 / 2^16 / 2^17, one graph")
    specs = [(150, 0.002, 32768, True), (180, 0.003, 65536, False),
             (120, 0.0035, 131072, True)]
    ctxs = []
    for i, (side, sig, nf, binned) in enumerate(specs):
        inp = workload.make_inputs(npix_side=side, sigma=sig, seed=60 + i)
        ctxs.append(workload.build_context(inp, nf, 512, binned=binned))
    ctxs[1].set_priors(sigma_width=0.1)
    W = 24
    th = synth.walker_thetas(W, seed=5)
    stride = th.shape[1]
    d_th = hiprt.DeviceBuffer(th.nbytes)
    d_ll = hiprt.DeviceBuffer(W * 8)
    d_st = hiprt.DeviceBuffer(W * 4)
    stream = hiprt.Stream()
    d_th.upload(th)
    g = _lib.FieldGraph(ctxs, d_th.value, W, stride, d_ll.value, d_st.value,
                        stream.ptr.value)
    for r in range(4):
        t = synth.walker_thetas(W, seed=100 + r)
        d_th.upload(t)
        g.launch(stream.ptr.value)
        stream.synchronize()
        ll = np.zeros(W)
        st = np.zeros(W, dtype=np.int32)
        d_ll.download(ll)
        d_st.download(st)
        want = np.zeros(W)
        for c in ctxs:                       # eager, between the replays
            v, s = c.loglike_batch(t)
            assert np.all(s == 0)
            want = want + v
        assert np.all(st == 0)
        np.testing.assert_allclose(ll, want, rtol=0, atol=1e-9)
        _say("replay %d matches" % r)


def test_field_graph_refuses_after_workspace_growth():
    """A larger eager batch replaces a context's workspace; the graph that
    captured the old one must refuse to launch (POFD_E_ARG) instead of
    touching freed memory, and a graph captured again works."""
    _say("graph invalidated by a larger eager batch")
    inp = workload.make_inputs(npix_side=120, sigma=0.003, seed=70)
    ctx = workload.build_context(inp, 32768, 512, binned=True)
    W = 8
    th = synth.walker_thetas(W, seed=9)
    stride = th.shape[1]
    d_th = hiprt.DeviceBuffer(th.nbytes)
    d_ll = hiprt.DeviceBuffer(W * 8)
    d_st = hiprt.DeviceBuffer(W * 4)
    stream = hiprt.Stream()
    d_th.upload(th)
    g = _lib.FieldGraph([ctx], d_th.value, W, stride, d_ll.value,
                        d_st.value, stream.ptr.value)
    g.launch(stream.ptr.value)
    stream.synchronize()
    want, s = ctx.loglike_batch(th)
    ll = np.zeros(W)
    d_ll.download(ll)
    np.testing.assert_array_equal(ll, want)
    big = synth.walker_thetas(4 * W, seed=10)
    ctx.loglike_batch(big)                 # grows the workspace
    with pytest.raises(_lib.PofdError) as ei:
        g.launch(stream.ptr.value)
    assert ei.value.code == 5  # POFD_E_ARG
    g2 = _lib.FieldGraph([ctx], d_th.value, W, stride, d_ll.value,
                         d_st.value, stream.ptr.value)
    g2.launch(stream.ptr.value)
    stream.synchronize()
    d_ll.download(ll)
    np.testing.assert_array_equal(ll, want)
