"""The measured configurations not covered elsewhere, against the oracle.

* The small CPU-runnable fit: a 720x720 unbinned map, nflux 2^17, one
  likelihood at the generating theta and a 40-walker x 200-step stretch-move
  chain whose stored log likelihoods are the direct ones.
* nflux 2^18 (the largest transform): 16 walkers, binned and unbinned.
* Uneven fields (2e5-4e6 pixels, nflux 2^15-2^18, noise 1.5-4 mJy), each
  field's batch against the oracle.
* Delayed acceptance at the headline size: 512 walkers at 2^17 with an
  8 -> 3x256 -> 1 surrogate trained on the chain's own pairs; stored log
  likelihoods exact, fewer exact evaluations than proposals, knot posterior
  medians (4000 steps after a 2000-step burn-in, the first 500 of each
  chain dropped) within 0.1 sigma of a plain chain's.

Exact parity (1e-6 against the extended-precision oracle) goes through
oracle_side.check_exact_parity, which also settles the rounding-decided
split index of the reference's argmin.
"""
import numpy as np
import pytest

from oracle_side import build_oracle, check_exact_parity
from pofd_affine_amd import _lib, synth, workload

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-config-case] " + msg, flush=True)


def test_small_fit_getlike_and_chain():
    _say("720x720 unbinned, nflux 2^17: getLike at the true theta")
    inp = workload.make_inputs(npix_side=720, seed=720)
    ctx = workload.build_context(inp, 131072, 2048, binned=False)
    t0 = synth.true_theta()[None, :].astype(np.float32)
    ll, st = ctx.loglike_batch(t0)
    assert st[0] == 0
    L = build_oracle(inp, 131072, 2048, False, t0[0])
    check_exact_parity(ctx, L, t0, ll, max_off=1)
    _say("40 walkers x 200 steps")
    nw = 40
    th = synth.walker_thetas(nw, seed=720)
    npar = th.shape[1]
    state = np.zeros(npar, dtype=np.uint8)
    state[8] = 1
    state[9:] = 2
    ens = _lib.Ensemble(ctx, _lib.TARGET_1D, nw, npar, seed=5)
    ens.set_param_state(state)
    ens.set_exponent_mode(True)
    ens.set_positions(th)
    ens.run(200, thin=1)
    cur, llc, nacc, nerr = ens.state()
    ref, st2 = ctx.loglike_batch(cur)
    assert np.all(st2 == 0)
    np.testing.assert_array_equal(llc, ref)
    assert np.all(nerr == 0)
    acc = nacc.sum() / (nw * 200.0)
    _say("acceptance %.3f" % acc)
    assert 0.05 < acc < 0.9
    ch, cll = ens.chain()
    assert ch.shape == (200, nw, npar) and np.all(np.isfinite(cll))
    # a walker that moved carries the flux and flux^2 integrals of its
    # position in the bonus parameters (pofdMCMC::fillBonusParams)
    import pyoracle as O
    moved = np.nonzero(nacc > 0)[0]
    assert moved.size > nw // 2
    for w in moved[:4]:
        m = O.Knots(inp["knots"], cur[w][:8])
        assert cur[w][9] == pytest.approx(m.splineint(1), rel=1e-6)
        assert cur[w][10] == pytest.approx(m.splineint(2), rel=1e-6)


@pytest.mark.parametrize("binned", [True, False])
def test_nflux_2p18(binned):
    _say("nflux 2^18, 16 walkers, binned=%s" % binned)
    inp = workload.make_inputs(npix_side=500, seed=18)
    ctx = workload.build_context(inp, 262144, 2048, binned=binned)
    th = synth.walker_thetas(16, seed=18)
    ll, st = ctx.loglike_batch(th)
    assert np.all(st == 0)
    L = build_oracle(inp, 262144, 2048, binned, th[0])
    check_exact_parity(ctx, L, th, ll, max_off=2)


def test_uneven_fields_against_oracle():
    """Each of the eight uneven fields (one context each, as the captured
    multi-field graph runs them) against the oracle for 8 walkers."""
    th = synth.walker_thetas(8, seed=77)
    for side, sigma, nflux, seed in workload.uneven_fields():
        _say("field %dx%d, sigma %.2f mJy, nflux %d" % (side, side,
                                                        1e3 * sigma, nflux))
        inp = workload.make_inputs(npix_side=side, sigma=sigma, seed=seed)
        ctx = workload.build_context(inp, nflux, 2048, binned=False)
        ll, st = ctx.loglike_batch(th)
        assert np.all(st == 0)
        L = build_oracle(inp, nflux, 2048, False, th[0])
        check_exact_parity(ctx, L, th, ll, max_off=2)


def test_delayed_acceptance_headline():
    _say("DA: 512 walkers, nflux 2^17, binned 1000x1000")
    inp = workload.make_inputs(npix_side=1000, seed=42)
    ctx = workload.build_context(inp, 131072, 2048, binned=True)
    W, nk = 512, 8
    th0 = synth.walker_thetas(W, seed=777, scatter=0.01)
    npar = th0.shape[1]
    state = np.zeros(npar, dtype=np.uint8)
    state[nk] = 1
    state[nk + 1:] = 2

    def make(pos, seed):
        e = _lib.Ensemble(ctx, _lib.TARGET_1D, W, npar, seed=seed)
        e.set_param_state(state)
        e.set_exponent_mode(True)
        e.set_positions(pos)
        return e

    ens = make(th0, 31)
    ens.run(2000, thin=0)  # from a 0.01 scatter to the posterior's width
    ens.run(100, thin=1)
    ch, cll = ens.chain()
    start, _, _, _ = ens.state()
    net = _lib.Mlp(nk, 256, seed=5)
    tr, ho = net.train(ch[:, :, :nk].reshape(-1, nk), cll.reshape(-1),
                       epochs=60, batch=512, lr=1e-3, holdout_frac=0.1,
                       seed=6)
    _say("surrogate holdout RMS %.3f" % ho)
    assert ho <= 0.5
    steps = 4000
    chains = {}
    for name, use_da, seed in (("plain", False, 41), ("da", True, 41),
                               ("control", False, 43)):
        e = make(start, seed)
        if use_da:
            assert e.set_surrogate(net, list(range(nk)))
        n0 = e.counts()
        e.run(steps, thin=1)
        n1 = e.counts()
        c, _ = e.chain()
        c = c[500:]  # the chains' common start is not yet a posterior draw
        cur, llc, _, _ = e.state()
        ref, st = ctx.loglike_batch(cur)
        assert np.all(st == 0)
        np.testing.assert_array_equal(llc, ref)  # stored logL are exact
        chains[name] = (c[:, :, :nk].reshape(-1, nk).astype(np.float64),
                        n1[1] - n0[1], n1[0] - n0[0])
        _say("%s: %d exact evaluations of %d proposals" % (
            name, n1[1] - n0[1], n1[0] - n0[0]))
    flat_p = chains["plain"][0]
    sig = flat_p.std(axis=0)
    dmed = (np.median(chains["da"][0], axis=0) -
            np.median(flat_p, axis=0)) / sig
    dctl = (np.median(chains["control"][0], axis=0) -
            np.median(flat_p, axis=0)) / sig
    _say("median shifts DA %s control %s" % (np.round(dmed, 3),
                                             np.round(dctl, 3)))
    assert chains["da"][1] < chains["da"][2]   # exact fraction below 1
    assert np.max(np.abs(dmed)) < 0.1
