"""The log-likelihood surrogate (C ABI pofd_mlp_*) and delayed acceptance in
the stretch move (pofd_ens_set_surrogate).

The network's loss and gradient are checked against the CPU oracle
(oracle/mlp_ref.py, itself pinned by finite differences) with the GPU's
bfloat16 roundings emulated; training is checked by the held-out error it
reaches on a known function.  Delayed acceptance is checked by what the
mathematics fixes: with any surrogate, good or deliberately wrong, the chain
samples the exact target (detailed balance of the two-stage acceptance);
with a good one it spends fewer exact evaluations; the held-out gate
switches it off; and on the P(D) likelihood every stored log likelihood is
the exact one.
"""
import os
import sys

import numpy as np
import pytest

from pofd_affine_amd import _lib, synth, workload

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp_ref  # noqa: E402

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-surrogate-case] " + msg, flush=True)


def _random_net(nin, H, seed):
    net = _lib.Mlp(nin, H, seed=seed)
    rng = np.random.default_rng(seed)
    w = np.zeros(net.nparams, dtype=np.float32)
    sizes = [H * nin, H, H * H, H, H * H, H, H, 1]
    fan = [nin, 1, H, 1, H, 1, H, 1]
    for k, (o, sz) in enumerate(zip(net.offsets, sizes)):
        w[o:o + sz] = rng.normal(scale=np.sqrt(2.0 / fan[k]) if k % 2 == 0
                                 else 0.1, size=sz)
    net.set_params(w)
    return net, w


@pytest.mark.parametrize("nin,H,M", [(3, 16, 40), (8, 256, 300)])
def test_loss_and_gradient_match_oracle(nin, H, M):
    _say("gradient nin=%d H=%d M=%d" % (nin, H, M))
    net, w = _random_net(nin, H, seed=nin + H)
    rng = np.random.default_rng(7)
    xmean = rng.normal(size=nin).astype(np.float32)
    xstd = (0.5 + rng.random(nin)).astype(np.float32)
    ymean, ystd = -3.0, 2.5
    net.set_standardization(xmean, xstd, ymean, ystd)
    x = (xmean + xstd * rng.normal(size=(M, nin))).astype(np.float32)
    y = ymean + ystd * rng.normal(size=M)
    loss, g = net.loss_grad(x, y)
    xs = mlp_ref.standardize(x, xmean, xstd)
    ys = ((y - ymean) / ystd).astype(np.float32).astype(np.float64)
    lref, gref = mlp_ref.loss_grad(w, net.offsets, nin, H, xs, ys)
    assert loss == pytest.approx(lref, rel=1e-4)
    scale = np.max(np.abs(gref))
    np.testing.assert_allclose(g, gref, rtol=1e-3, atol=1e-3 * scale)
    # prediction in log-likelihood units
    out, _ = mlp_ref.forward(w, net.offsets, nin, H, xs)
    np.testing.assert_allclose(net.predict(x), out * ystd + ymean,
                               rtol=0, atol=1e-4 * ystd * np.max(np.abs(out)))
    # a full whitening matrix in place of the diagonal standardisation
    wmat = np.tril(rng.normal(size=(nin, nin))) + 2.0 * np.eye(nin)
    net.set_whitening(xmean, wmat, ymean, ystd)
    zs = mlp_ref.whiten(x, xmean, wmat)
    outw, _ = mlp_ref.forward(w, net.offsets, nin, H, zs)
    np.testing.assert_allclose(net.predict(x), outw * ystd + ymean, rtol=0,
                               atol=2e-3 * ystd * np.max(np.abs(outw)))
    net.set_standardization(xmean, xstd, ymean, ystd)
    # a plain gradient step lowers the loss
    net.set_params(w - 0.01 * g / scale)
    assert net.loss_grad(x, y)[0] < loss


def _quadratic(x, mean, sig):
    r = (x - mean) / sig
    return -0.5 * np.sum(r * r, axis=-1)


def test_training_reaches_small_holdout_error():
    _say("training on a quadratic")
    rng = np.random.default_rng(1)
    mean = np.array([1.0, -2.0, 0.5, 3.0])
    sig = np.array([0.5, 2.0, 0.1, 1.0])
    x = (mean + 2.0 * sig * rng.normal(size=(8192, 4))).astype(np.float32)
    y = _quadratic(x.astype(np.float64), mean, sig)
    net = _lib.Mlp(4, 128, seed=3)
    loss0 = net.loss_grad(x[:512], y[:512])
    tr, ho = net.train(x, y, epochs=150, batch=256, lr=2e-3,
                       holdout_frac=0.1, seed=5)
    _say("train RMS %.4f holdout RMS %.4f (std y %.2f)" % (tr, ho, y.std()))
    assert ho < 0.05 * y.std() and ho < 0.5
    assert tr < 0.05 * y.std()
    pred = net.predict(x[:512])
    assert np.sqrt(np.mean((pred - y[:512]) ** 2)) < 0.05 * y.std()
    assert loss0[0] > 0.5  # it did start untrained


def test_training_whitens_correlated_inputs():
    """Inputs correlated at 0.999 whose target depends on their small
    difference: the training's whitening (Cholesky of the covariance) puts
    that direction at unit scale before the bfloat16 rounding."""
    _say("training on correlated inputs")
    rng = np.random.default_rng(8)
    n = 8192
    u = rng.normal(size=n)
    v = rng.normal(size=n)
    a = 5.0 + u
    b = 5.0 + u + 0.0447 * v          # corr(a, b) = 0.999
    c = rng.normal(size=n)
    x = np.stack([a, b, c], axis=1)
    r = (b - a) / 0.0447               # the narrow direction, unit width
    y = -0.5 * (r * r + c * c + u * u)
    net = _lib.Mlp(3, 128, seed=4)
    tr, ho = net.train(x.astype(np.float32), y, epochs=120, batch=256,
                       lr=2e-3, holdout_frac=0.1, seed=9)
    _say("train RMS %.4f holdout RMS %.4f (std y %.2f)" % (tr, ho, y.std()))
    assert ho < 0.05 * y.std() and ho < 0.1


def _gauss_da(seed, surrogate_mean_shift, surrogate_width, nwalk=64):
    """An ensemble on a 6-parameter Gaussian (one fixed, one bonus) with a
    surrogate trained on a possibly wrong Gaussian."""
    mean = np.array([1.0, -2.0, 0.5, 3.0, 7.0, 0.0])
    sig = np.array([0.5, 2.0, 0.1, 1.0, 1.0, 1.0])
    ens = _lib.Ensemble(None, _lib.TARGET_GAUSS, nwalk, 6, scalefac=2.0,
                        seed=seed)
    ens.set_gauss_target(mean, sig)
    ens.set_param_state([0, 0, 0, 0, 1, 2])
    ens.set_exponent_mode(True)
    rng = np.random.default_rng(seed)
    xtr = (mean[:4] + 2.0 * sig[:4] * rng.normal(size=(8192, 4)))
    smean = mean[:4] + surrogate_mean_shift * sig[:4]
    ytr = _quadratic(xtr, smean, sig[:4] * surrogate_width)
    net = _lib.Mlp(4, 64, seed=seed)
    net.train(xtr.astype(np.float32), ytr, epochs=150, batch=256, lr=2e-3,
              holdout_frac=0.1, seed=seed)
    th = mean + 0.1 * sig * rng.normal(size=(nwalk, 6))
    th[:, 4] = 7.0
    ens.set_positions(th.astype(np.float32))
    return ens, net, mean, sig


@pytest.mark.parametrize("shift,width", [(0.0, 1.0), (0.7, 1.5)])
def test_delayed_acceptance_keeps_exact_target(shift, width):
    """Detailed balance of the two-stage acceptance: the chain samples the
    exact Gaussian whether the surrogate is right (shift 0, width 1) or
    wrong by 0.7 sigma and 50% in width."""
    _say("DA Gaussian shift=%.1f width=%.1f" % (shift, width))
    ens, net, mean, sig = _gauss_da(21, shift, width)
    assert ens.set_surrogate(net, [0, 1, 2, 3], gate=np.inf,
                             holdout_rms=0.0)
    ens.run(500, thin=0)
    ens.run(4000, thin=1)
    ch, ll = ens.chain()
    _, _, nacc, nerr = ens.state()
    nprop, nexact = ens.counts()
    _say("proposals %d exact %d accepted %d" % (nprop, nexact, nacc.sum()))
    assert nprop == 64 * 4500
    assert nexact < nprop
    assert nacc.sum() <= nexact
    assert np.all(ch[:, :, 4] == 7.0) and np.all(np.isnan(ch[:, :, 5]))
    r = (ch[..., :4].astype(np.float64) - mean[:4]) / sig[:4]
    np.testing.assert_allclose(ll, -0.5 * np.sum(r * r, axis=-1),
                               rtol=1e-12, atol=1e-12)
    for i in range(4):
        # standard error of the mean from 20 batch means of 200 steps (the
        # autocorrelation time here is 20-100 steps)
        bm = ch[:, :, i].astype(np.float64).mean(axis=1).reshape(20, 200)
        bm = bm.mean(axis=1)
        se = bm.std(ddof=1) / np.sqrt(20)
        m = bm.mean()
        assert abs(m - mean[i]) < 5 * se, (i, m, se)
        assert np.std(ch[:, :, i]) == pytest.approx(sig[i], rel=0.06), i


@pytest.mark.parametrize("shift,width", [(0.0, 1.0), (0.7, 1.5)])
def test_surrogate_transition_keeps_exact_target(shift, width):
    """The surrogate transition (8 surrogate stretch moves per walker and
    half-step, one exact correction at the end) samples the exact Gaussian
    whether the surrogate is right or 0.7 sigma and 50% in width wrong; the
    stored log likelihoods are the exact ones, and with a right surrogate
    the chain decorrelates faster than the plain one per step."""
    _say("surrogate transition Gaussian shift=%.1f width=%.1f" % (shift,
                                                                   width))
    ens, net, mean, sig = _gauss_da(23, shift, width)
    assert ens.set_surrogate(net, [0, 1, 2, 3], gate=np.inf,
                             holdout_rms=0.0)
    ens.set_surrogate_steps(8)
    ens.run(500, thin=0)
    n0 = ens.counts()
    ens.run(4000, thin=1)
    n1 = ens.counts()
    ch, ll = ens.chain()
    _, _, nacc, nerr = ens.state()
    _say("proposals %d exact %d" % (n1[0] - n0[0], n1[1] - n0[1]))
    assert n1[0] - n0[0] == 64 * 4000
    assert n1[1] - n0[1] <= n1[0] - n0[0]
    assert np.all(ch[:, :, 4] == 7.0) and np.all(np.isnan(ch[:, :, 5]))
    r = (ch[..., :4].astype(np.float64) - mean[:4]) / sig[:4]
    np.testing.assert_allclose(ll, -0.5 * np.sum(r * r, axis=-1),
                               rtol=1e-12, atol=1e-12)
    for i in range(4):
        bm = ch[:, :, i].astype(np.float64).mean(axis=1).reshape(20, 200)
        bm = bm.mean(axis=1)
        se = bm.std(ddof=1) / np.sqrt(20)
        m = bm.mean()
        assert abs(m - mean[i]) < 5 * se, (i, m, se)
        assert np.std(ch[:, :, i]) == pytest.approx(sig[i], rel=0.06), i
    if shift == 0.0:
        tau_st = np.mean([ens.acor(i)[0] for i in range(4)])
        ref, _, _, _ = _gauss_da(23, shift, width)
        ref.run(500, thin=0)
        ref.run(4000, thin=1)
        tau_plain = np.mean([ref.acor(i)[0] for i in range(4)])
        _say("tau plain %.1f surrogate transition %.1f" % (tau_plain,
                                                           tau_st))
        assert tau_st < 0.5 * tau_plain


def test_surrogate_transition_steps_limits():
    ens, net, _, _ = _gauss_da(7, 0.0, 1.0)
    for bad in (0, 65):
        with pytest.raises(_lib.PofdError):
            ens.set_surrogate_steps(bad)
    ens.set_surrogate_steps(64)
    ens.set_surrogate_steps(1)


def test_good_surrogate_saves_exact_evaluations():
    _say("DA savings")
    ens, net, mean, sig = _gauss_da(5, 0.0, 1.0)
    assert net.holdout_rms < 0.5
    assert ens.set_surrogate(net, [0, 1, 2, 3])  # default gate 0.5
    ens.run(1000, thin=0)
    nprop, nexact = ens.counts()
    _, _, nacc, _ = ens.state()
    # exact evaluations only for the first-stage survivors; nearly all of
    # them are accepted by the second stage
    assert nexact < 0.7 * nprop
    assert nacc.sum() > 0.8 * nexact


def test_gate_switches_delayed_acceptance_off():
    _say("DA gate")
    ens, net, _, _ = _gauss_da(6, 0.0, 1.0)
    assert not ens.set_surrogate(net, [0, 1, 2, 3], gate=0.5,
                                 holdout_rms=0.6)
    ens.run(50, thin=0)
    nprop, nexact = ens.counts()
    assert nprop == nexact == 64 * 50
    with pytest.raises(_lib.PofdError):
        ens.set_surrogate(net, [0, 1, 2, 5], gate=np.inf, holdout_rms=0.0)
    with pytest.raises(_lib.PofdError):
        ens.set_surrogate(net, [0, 1, 2], gate=np.inf, holdout_rms=0.0)


def test_delayed_acceptance_on_pofd1d():
    """Burn-in on the exact 1D P(D) likelihood, a network trained on the
    (theta, logL) pairs it produced, then delayed acceptance: every
    walker's log likelihood is still the exact one at its position."""
    _say("DA on the 1D P(D) likelihood")
    inp = workload.make_inputs(npix_side=200, seed=7)
    ctx = workload.build_context(inp, 32768, 512, binned=True)
    nw = 64
    th = synth.walker_thetas(nw, seed=29)
    npar = th.shape[1]
    ens = _lib.Ensemble(ctx, _lib.TARGET_1D, nw, npar, seed=13)
    state = np.zeros(npar, dtype=np.uint8)
    state[8] = 1
    state[9:] = 2
    ens.set_param_state(state)
    ens.set_exponent_mode(True)
    ens.set_positions(th)
    ens.run(150, thin=1)
    ch, cll = ens.chain()
    x = ch[50:, :, :8].reshape(-1, 8)
    y = cll[50:].reshape(-1)
    net = _lib.Mlp(8, 128, seed=2)
    tr, ho = net.train(x, y, epochs=100, batch=256, lr=1e-3,
                       holdout_frac=0.1, seed=3)
    _say("train RMS %.3f holdout RMS %.3f" % (tr, ho))
    assert ens.set_surrogate(net, list(range(8)), gate=np.inf,
                             holdout_rms=0.0)
    n0 = ens.counts()
    ens.run(100, thin=1)
    n1 = ens.counts()
    cur, ll, nacc, nerr = ens.state()
    ref, st = ctx.loglike_batch(cur)
    assert np.all(st == 0)
    np.testing.assert_array_equal(ll, ref)
    _say("exact %d of %d proposals" % (n1[1] - n0[1], n1[0] - n0[0]))
    assert n1[1] - n0[1] < n1[0] - n0[0]
    ch2, cll2 = ens.chain()
    assert ch2.shape[0] == 250 and np.all(np.isfinite(cll2))


def test_surrogate_transition_on_pofd1d():
    """The surrogate transition on the 1D P(D) likelihood: every walker's
    stored log likelihood is the exact one at its position, and at most one
    exact evaluation per walker and half-step."""
    _say("surrogate transition on the 1D P(D) likelihood")
    inp = workload.make_inputs(npix_side=200, seed=7)
    ctx = workload.build_context(inp, 32768, 512, binned=True)
    nw = 64
    th = synth.walker_thetas(nw, seed=31)
    npar = th.shape[1]
    ens = _lib.Ensemble(ctx, _lib.TARGET_1D, nw, npar, seed=17)
    state = np.zeros(npar, dtype=np.uint8)
    state[8] = 1
    state[9:] = 2
    ens.set_param_state(state)
    ens.set_exponent_mode(True)
    ens.set_positions(th)
    ens.run(150, thin=1)
    ch, cll = ens.chain()
    net = _lib.Mlp(8, 128, seed=2)
    tr, ho = net.train(ch[50:, :, :8].reshape(-1, 8), cll[50:].reshape(-1),
                       epochs=100, batch=256, lr=1e-3, holdout_frac=0.1,
                       seed=3)
    _say("train RMS %.3f holdout RMS %.3f" % (tr, ho))
    assert ens.set_surrogate(net, list(range(8)), gate=np.inf,
                             holdout_rms=0.0)
    ens.set_surrogate_steps(6)
    n0 = ens.counts()
    _, _, a0, _ = ens.state()
    ens.run(60, thin=1)
    n1 = ens.counts()
    cur, ll, a1, nerr = ens.state()
    ref, st = ctx.loglike_batch(cur)
    assert np.all(st == 0)
    np.testing.assert_array_equal(ll, ref)
    _say("exact %d of %d walker half-steps, accepted %d" % (
        n1[1] - n0[1], n1[0] - n0[0], (a1 - a0).sum()))
    assert 0 < n1[1] - n0[1] <= n1[0] - n0[0]
    assert (a1 - a0).sum() > 0
    _, cll2 = ens.chain()
    assert np.all(np.isfinite(cll2))


def test_delayed_acceptance_on_pofd2d():
    """Delayed acceptance on the two-band likelihood: a 12-input network
    (band-1 knots, colour sigma and offset knots) trained on the chain's own
    pairs; afterwards every walker's log likelihood is still the exact one
    and fewer exact evaluations were spent than proposals made."""
    _say("DA on the 2D P(D) likelihood")
    from pofd_affine_amd import synth2d, workload2d
    from oracle2d_side import make_inputs2d
    inp = make_inputs2d(npix_side=120, seed=5, nbins=60)
    ctx = workload2d.build_context2d(inp, 512, synth2d.true_theta(),
                                     nedge=64, binned=True)
    nw = 32
    th = synth2d.walker_thetas(nw, seed=41, scatter=0.01)
    npar = th.shape[1]
    nfit = 12
    state = np.zeros(npar, dtype=np.uint8)
    state[nfit:nfit + 2] = 1      # sigma multipliers fixed
    state[nfit + 2:] = 2          # bonus
    ens = _lib.Ensemble(ctx, _lib.TARGET_2D, nw, npar, seed=19)
    ens.set_param_state(state)
    ens.set_exponent_mode(True)
    ens.set_positions(th)
    ens.run(160, thin=1)
    ch, cll = ens.chain()
    x = ch[40:, :, :nfit].reshape(-1, nfit)
    y = cll[40:].reshape(-1)
    net = _lib.Mlp(nfit, 128, seed=7)
    tr, ho = net.train(x, y, epochs=150, batch=128, lr=1e-3,
                       holdout_frac=0.1, seed=8)
    _say("2D surrogate train RMS %.3f holdout RMS %.3f" % (tr, ho))
    assert ens.set_surrogate(net, list(range(nfit)), gate=np.inf,
                             holdout_rms=0.0)
    n0 = ens.counts()
    ens.run(60, thin=1)
    n1 = ens.counts()
    cur, ll, nacc, nerr = ens.state()
    ref, st = ctx.loglike2d_batch(cur)
    assert np.all(st == 0)
    np.testing.assert_allclose(ll, ref, rtol=0, atol=1e-9)
    _say("2D exact %d of %d proposals" % (n1[1] - n0[1], n1[0] - n0[0]))
    assert n1[1] - n0[1] < n1[0] - n0[0]
    assert np.all(ch[:, :, nfit:nfit + 2] == th[0, nfit:nfit + 2])
