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
