"""Throughput of the batched 1D P(D) log likelihood on one MI355X.

One step = one exact log-likelihood evaluation of a batch of walkers
(calcLike::getLogLike for every walker, src/calcLike.cc:884-955) at
nflux = 2^17 on a synthetic 1000x1000-pixel map histogrammed into 1000 bins,
with theta already resident in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("P(D) logL evals/s (2^17, 256 walkers); DA-vs-plain ESS/s, "
          "|ΔlogL|≤1e-6")
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of each kernel at the default workload (256 walkers,
# nflux 2^17, binned), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs
# scaled by a plain copy of known size (fetched bytes = 2.0 x FETCH_SIZE KiB,
# written = 1.0 x WRITE_SIZE KiB); None where not profiled or for another
# workload.
TRAFFIC = {
    "prep": (2.0 * 2098.1 + 20512.0) * 1024,
    "forward": (2.0 * 19755.7 + 35308.2) * 1024,
    "inverse": (2.0 * 2774.5 + 279619.0) * 1024,
    "finalize": (2.0 * 10061.2 + 64.0) * 1024,
}
FP64_PEAK_TFLOPS = 78.6
# FP64 FLOP per launch of each kernel at the default workload, from
# rocprofv3 --pmc SQ_INSTS_VALU_{FMA,ADD,MUL}_F64 and
# SQ_INSTS_VALU_MFMA_MOPS_F64: 64 x (2 FMA + ADD + MUL) + 512 x MOPS (a
# 16x16x4 f64 MFMA is 2048 FLOP, counted as 4 MOPS).  Divided by the live
# kernel time for the FP64 roofline.
FP64_FLOP = {
    "prep": 2.738e8,
    "forward": 2.924e9,
    "inverse": 1.943e9,
    "finalize": 6.39e7,
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--walkers", type=int, default=256)
    ap.add_argument("--nflux", type=int, default=131072)
    ap.add_argument("--ninterp", type=int, default=2048)
    ap.add_argument("--side", type=int, default=1000)
    ap.add_argument("--unbinned-steps", type=int, default=10)
    ap.add_argument("--cpu-evals", type=int, default=1024)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pcie-steps", type=int, default=10,
                    help="steps through the host-buffer entry point")
    ap.add_argument("--steps-2d", type=int, default=10,
                    help="timed 2D batches (0: skip the 2D leg)")
    ap.add_argument("--walkers-2d", type=int, default=32)
    ap.add_argument("--n-2d", type=int, default=2048)
    ap.add_argument("--cpu-evals-2d", type=int, default=16)
    ap.add_argument("--only-2d", action="store_true",
                    help="run the 2D leg alone (for profiling it)")
    ap.add_argument("--da-steps", type=int, default=2000,
                    help="steps of each of the plain and the delayed-"
                         "acceptance chains compared (0 skips that leg)")
    ap.add_argument("--da-walkers", type=int, default=512)
    ap.add_argument("--da-burn", type=int, default=500)
    ap.add_argument("--da-train-steps", type=int, default=100)
    ap.add_argument("--da-epochs", type=int, default=60)
    ap.add_argument("--st-steps", type=int, default=8,
                    help="surrogate steps per half-step of the surrogate-"
                         "transition chain of the delayed-acceptance leg "
                         "(0 or 1 skips that chain)")
    ap.add_argument("--only-da", action="store_true",
                    help="run the delayed-acceptance leg only")
    ap.add_argument("--da2d-steps", type=int, default=500,
                    help="steps of each 2D chain of the two-band delayed-"
                         "acceptance leg (0 skips that leg)")
    ap.add_argument("--only-da2d", action="store_true",
                    help="run the two-band delayed-acceptance leg only")
    ap.add_argument("--multifield-steps", type=int, default=20,
                    help="steps of the uneven multi-field graph leg (0 skips)")
    ap.add_argument("--only-multifield", action="store_true")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="BYTES",
                    help="first run a plain device copy of BYTES (calibrates "
                         "the rocprofv3 FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the walker batch "
              "on a single device (multi-GPU sharding is a design note, not "
              "an executable path); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from pofd_affine_amd import hiprt, synth, workload

    if args.copy_calib > 0:
        import ctypes
        calib = ctypes.CDLL(os.path.join(ROOT, "pofd_affine_amd", "_build",
                                         "libpofd_calib.so"))
        calib.pofd_calib_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_uint64, ctypes.c_void_p]
        nb = (args.copy_calib // 16) * 16
        src, dst = hiprt.DeviceBuffer(nb), hiprt.DeviceBuffer(nb)
        for _ in range(3):
            if calib.pofd_calib_copy(src.value, dst.value, nb // 8, None):
                raise RuntimeError("calibration copy failed")
        hiprt.device_synchronize()
        del src, dst

    if args.only_2d:
        print(json.dumps({"twod": bench_2d(args, hiprt)}))
        return

    if args.only_da2d:
        print(json.dumps({"da2d": bench_da2d(args)}))
        return
    W, N = args.walkers, args.nflux
    inp = workload.make_inputs(npix_side=args.side, seed=42)
    if args.only_da:
        print(json.dumps({"da": bench_da(args, inp)}))
        return
    if args.only_multifield:
        print(json.dumps({"multifield": bench_multifield(args, hiprt)}))
        return
    th = synth.walker_thetas(W, seed=1234)
    stride = th.shape[1]

    def run(binned, steps, warmup):
        ctx = workload.build_context(inp, N, args.ninterp, binned=binned)
        ctx.set_stage_timing(False)
        d_th = hiprt.DeviceBuffer(th.nbytes)
        d_th.upload(th)
        d_ll = hiprt.DeviceBuffer(W * 8)
        d_st = hiprt.DeviceBuffer(W * 4)
        stream = hiprt.Stream()
        for _ in range(warmup):
            ctx.loglike_batch_device(d_th.value, W, stride, d_ll.value,
                                     d_st.value, stream.ptr.value)
        stream.synchronize()
        hiprt.device_synchronize()
        evs = [hiprt.Event() for _ in range(steps + 1)]
        t0 = time.perf_counter()
        evs[0].record(stream)
        for i in range(steps):
            ctx.loglike_batch_device(d_th.value, W, stride, d_ll.value,
                                     d_st.value, stream.ptr.value)
            evs[i + 1].record(stream)
        stream.synchronize()
        hiprt.device_synchronize()
        wall = time.perf_counter() - t0
        per = [evs[i].elapsed_ms(evs[i + 1]) for i in range(steps)]
        # per-kernel durations from further steps with HIP events around
        # every kernel stage (kept out of the timed steps: the events add
        # a few microseconds per stage boundary)
        ctx.set_stage_timing(True)
        for _ in range(max(5, steps // 5)):
            ctx.loglike_batch_device(d_th.value, W, stride, d_ll.value,
                                     d_st.value, stream.ptr.value)
        stream.synchronize()
        stage_ms, nl = ctx.stage_times()
        stage_ms = {k: v / max(nl, 1) for k, v in stage_ms.items()}
        ctx.set_stage_timing(False)
        ll = np.zeros(W)
        st = np.zeros(W, dtype=np.int32)
        d_ll.download(ll)
        d_st.download(st)
        return dict(wall=wall, per_ms=per, ll=ll, st=st, ctx=ctx,
                    stage_ms=stage_ms)

    b = run(True, args.steps, args.warmup)
    ms_step = 1e3 * b["wall"] / args.steps
    evals_per_step = W
    value = evals_per_step / (b["wall"] / args.steps)
    batch_ms = float(np.mean(b["per_ms"]))
    nbins = inp["nbins"]
    # Algorithmic HBM bytes of one evaluation: the three stage boundaries of
    # the transform pipeline that need a global dependency -- R -> forward
    # transform, spectrum -> inverse transform, P(D) -> split/normalise/
    # logL -- each one write plus one read of 8N bytes, plus the binned data
    # amortised over the batch.  Per kernel: forward holds boundary 1 (write
    # and read) and writes boundary 2; inverse reads 2 and writes 3;
    # finalize reads 3 and the data.
    b_eval = 48.0 * N + 4.0 * nbins / W
    b_stage = {"forward": 24.0 * N, "inverse": 16.0 * N,
               "finalize": 8.0 * N + 4.0 * nbins / W}
    stage_ms = b["stage_ms"]
    dom = max(b_stage, key=lambda k: stage_ms[k])
    kern_ms = stage_ms[dom]
    achieved = b_stage[dom] * W / (kern_ms * 1e-3) / 1e9
    pipe_achieved = b_eval * W / (batch_ms * 1e-3) / 1e9
    default_workload = (N == 131072 and W == 256 and args.side == 1000 and
                        args.ninterp == 2048)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "1D P(D) log likelihood, nflux=%d, ninterp=%d, "
                        "%d walkers per launch, %dx%d-pixel synthetic "
                        "map binned into %d bins, 120-bin Gaussian beam "
                        "histogram, sigma 2.5 mJy" % (N, args.ninterp, W,
                                                      args.side, args.side,
                                                      nbins),
            "nflux": N, "walkers": W, "map_side": args.side,
            "nbins": nbins, "ninterp": args.ninterp,
        },
        "status_ok": int(np.sum(b["st"] == 0)),
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": TRAFFIC.get(dom) if default_workload else None,
            "kernel": dom + "1d_kernel",
            "kernel_ms": kern_ms,
            "bytes_alg_per_launch": b_stage[dom] * W,
            "stage_ms": stage_ms,
            "pipeline": {
                "achieved": pipe_achieved,
                "frac": pipe_achieved / HBM_PEAK_GBS,
                "batch_ms": batch_ms,
                "bytes_alg_per_eval": b_eval,
            },
        },
    }
    if default_workload:
        # the same live kernel times against the FP64 peak, with the FLOP
        # per launch the SQ counters give for this workload
        fp = {k: FP64_FLOP[k] / (stage_ms[k] * 1e-3) / 1e12
              for k in FP64_FLOP if stage_ms.get(k, 0) > 0}
        out["roofline"]["fp64"] = {
            "kernel": dom + "1d_kernel", "achieved": fp[dom],
            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": fp[dom] / FP64_PEAK_TFLOPS,
            "flop_per_launch": FP64_FLOP[dom],
            "per_kernel": {k + "1d_kernel": {"achieved": v,
                                             "frac": v / FP64_PEAK_TFLOPS}
                           for k, v in fp.items()},
        }
    if args.unbinned_steps > 0:
        u = run(False, args.unbinned_steps, 2)
        out["unbinned"] = {
            "value": W / (u["wall"] / args.unbinned_steps),
            "unit": "evals/s",
            "batch_ms": float(np.mean(u["per_ms"])),
            "stage_ms": u["stage_ms"],
            "pixels": int(inp["data"].size),
            "status_ok": int(np.sum(u["st"] == 0)),
        }
    if args.pcie_steps > 0:
        # host-buffer entry point: theta in and logL/status out over PCIe
        ctx = b["ctx"]
        ctx.loglike_batch(th)
        t0 = time.perf_counter()
        for _ in range(args.pcie_steps):
            ctx.loglike_batch(th)
        dt = (time.perf_counter() - t0) / args.pcie_steps
        out["pcie_inclusive"] = {"value": W / dt, "unit": "evals/s",
                                 "ms_per_step": 1e3 * dt}
    if args.copy_calib > 0:
        out["copy_calib_bytes"] = (args.copy_calib // 16) * 16
    if args.steps_2d > 0:
        out["twod"] = bench_2d(args, hiprt)
    if args.da_steps > 0:
        out["da"] = bench_da(args, inp)
    if args.da2d_steps > 0 and args.steps_2d > 0:
        out["da2d"] = bench_da2d(args)
    if args.multifield_steps > 0:
        out["multifield"] = bench_multifield(args, hiprt)
    if not args.no_cpu and args.cpu_evals > 0:
        out["cpu_baseline"], par = cpu_baseline_1d(args, inp, th, True,
                                                   b["ll"], b["st"], b["ctx"])
        out.update(par)
        if "unbinned" in out:
            ub, upar = cpu_baseline_1d(args, inp, th, False, u["ll"],
                                       u["st"], u["ctx"],
                                       evals=args.cpu_evals // 2)
            out["unbinned"]["cpu_baseline"] = ub
            out["unbinned"].update(upar)
    print(json.dumps(out))


def host_cpu():
    """(usable host threads, capped at the 16 a one-GPU box grants; CPU
    model name)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return min(16, n), model


def cpu_baseline_1d(args, inp, th, binned, gpu_ll, gpu_st, ctx, evals=None):
    """The oracle (a plain-double restatement of the reference CPU path) on
    the host: one evaluation per worker thread at a time, as the
    reference's MPI slaves do (the master computes nothing,
    src/pofdMCMC.cc:200), with nproc - 1 and nproc workers, the per-stage
    split in the reference's TIMING categories, and the GPU's distance from
    the double oracle and from the extended-precision (exact-math) one --
    also at the device's own split index for the walkers whose split
    search landed on the other of two values that agree to within double
    rounding (the argmin of src/PDFactory.cc:783-795)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_side import build_oracle
    N, W = args.nflux, len(th)
    nproc, model = host_cpu()
    if args.cpu_threads:
        nproc = args.cpu_threads
    ne = evals or args.cpu_evals
    reps = -(-ne // W)
    sample = np.concatenate([th] * reps)[:ne]
    L = build_oracle(inp, N, args.ninterp, binned, th[0])
    L.loglike_batch(th[:nproc], nproc)  # one warm-up evaluation per worker
    runs = {}
    for nw in sorted({max(1, nproc - 1), nproc}):
        L.set_timing(True)
        t0 = time.perf_counter()
        L.loglike_batch(sample, nw)
        dt = time.perf_counter() - t0
        tm = L.timing()
        tot = sum(tm.values())
        runs[nw] = dict(value=ne / dt, wall_s=dt, cpu_s=tot,
                        stage_frac={k: v / tot for k, v in tm.items()})
    L.set_timing(False)
    best = runs[nproc]
    cb = {
        "value": best["value"], "unit": "evals/s", "cores": nproc,
        "kind": "port",
        "sample": "%d evaluations (the %d walkers of the same workload, "
                  "%s, nflux %d, cycled) per worker count, one evaluation "
                  "per host thread at a time; own radix-2 FFT (FFTW is "
                  "absent)" % (ne, W, "binned" if binned else "unbinned", N),
        "cpu_model": model,
        "workers": {str(k): {"value": v["value"], "wall_s": v["wall_s"]}
                    for k, v in runs.items()},
        "stage_frac": best["stage_frac"],
        "cpu_s_per_eval": best["cpu_s"] / ne,
    }
    dbl, dst = L.loglike_batch(th, nproc)
    L.set_precise(1)
    ext, est = L.loglike_batch(th, nproc)
    ok = (dst == 0) & (est == 0) & (gpu_st == 0)
    par = {}
    if ok.any():
        at_dev = ext.copy()
        amb = []
        for w in np.nonzero(ok & (np.abs(ext - gpu_ll) > 1e-6))[0]:
            w = int(w)
            gsplit = int(ctx.get_debug(w)["split"])
            L.initpd(0, th[w])
            esplit = L.getpd(0, 0, float(th[w][args_sigmult_index(L)]),
                             setlog=False)["split"]
            if gsplit != esplit:
                L.set_split_shift(gsplit - esplit)
                v, st1 = L.loglike_batch(th[w:w + 1], 1)
                L.set_split_shift(0)
                if st1[0] == 0:
                    at_dev[w] = v[0]
                amb.append([w, gsplit, int(esplit)])
        par = {
            "parity_max_abs_dlogL": float(np.max(np.abs(dbl[ok] - gpu_ll[ok]))),
            "parity_max_abs_dlogL_vs_exact":
                float(np.max(np.abs(ext[ok] - gpu_ll[ok]))),
            "parity_max_abs_dlogL_vs_exact_at_device_split":
                float(np.max(np.abs(at_dev[ok] - gpu_ll[ok]))),
            "split_ambiguous_walkers": amb,
            "oracle_double_vs_exact": float(np.max(np.abs(ext[ok] - dbl[ok]))),
            "parity_walkers": int(ok.sum()),
        }
    return cb, par


def args_sigmult_index(L):
    """Index of the sigma multiplier in theta: after the knots."""
    return L.nknots


def bench_multifield(args, hiprt):
    """Uneven fields sharing one theta: 8 synthetic maps of 2e5-4e6 pixels
    (log-uniform), noise 1.5-4 mJy, nflux 2^15-2^18 per field, unbinned,
    256 walkers; the eight fields' batches captured as one HIP graph
    (pofd1d_graph_create) whose launch gives the summed log likelihood.
    Per-step latency percentiles beside evals/s, and the same work launched
    eagerly field by field for comparison."""
    from pofd_affine_amd import _lib, synth, workload
    fields = workload.uneven_fields()
    nfield = len(fields)
    sides = np.array([f[0] for f in fields])
    sigmas = np.array([f[1] for f in fields])
    nfluxes = np.array([f[2] for f in fields])
    ctxs = []
    for side, sigma, nflux, seed in fields:
        inp = workload.make_inputs(npix_side=side, sigma=sigma, seed=seed)
        ctxs.append(workload.build_context(inp, nflux, args.ninterp, binned=False))
    W = args.walkers
    th = synth.walker_thetas(W, seed=77)
    stride = th.shape[1]
    d_th = hiprt.DeviceBuffer(th.nbytes)
    d_th.upload(th)
    d_ll = hiprt.DeviceBuffer(W * 8)
    d_st = hiprt.DeviceBuffer(W * 4)
    stream = hiprt.Stream()
    g = _lib.FieldGraph(ctxs, d_th.value, W, stride, d_ll.value, d_st.value,
                        stream.ptr.value)
    for _ in range(2):
        g.launch(stream.ptr.value)
    stream.synchronize()
    K = args.multifield_steps
    evs = [hiprt.Event() for _ in range(K + 1)]
    t0 = time.perf_counter()
    evs[0].record(stream)
    for i in range(K):
        g.launch(stream.ptr.value)
        evs[i + 1].record(stream)
    stream.synchronize()
    wall = time.perf_counter() - t0
    per = np.array([evs[i].elapsed_ms(evs[i + 1]) for i in range(K)])
    ll = np.zeros(W)
    st = np.zeros(W, dtype=np.int32)
    d_ll.download(ll)
    d_st.download(st)
    # eager: the same batches launched field by field
    outs = [(hiprt.DeviceBuffer(W * 8), hiprt.DeviceBuffer(W * 4))
            for _ in range(nfield)]

    def eager():
        for c, (a, b) in zip(ctxs, outs):
            c.loglike_batch_device(d_th.value, W, stride, a.value, b.value,
                                   stream.ptr.value)
    eager()
    stream.synchronize()
    eevs = [hiprt.Event() for _ in range(K + 1)]
    eevs[0].record(stream)
    for i in range(K):
        eager()
        eevs[i + 1].record(stream)
    stream.synchronize()
    eper = np.array([eevs[i].elapsed_ms(eevs[i + 1]) for i in range(K)])
    want = np.zeros(W)
    for a, _ in outs:
        v = np.zeros(W)
        a.download(v)
        want += v
    return {
        "workload": "%d unbinned fields sharing one theta, %d walkers, "
                    "pixels %s, sigma %s mJy, nflux %s; one captured HIP "
                    "graph per step" % (
                        nfield, W, [int(x) ** 2 for x in sides],
                        [round(1e3 * float(x), 2) for x in sigmas],
                        [int(x) for x in nfluxes]),
        "value": W / (wall / K), "unit": "evals/s", "steps": K,
        "step_ms_p50": float(np.percentile(per, 50)),
        "step_ms_p99": float(np.percentile(per, 99)),
        "eager_step_ms_p50": float(np.percentile(eper, 50)),
        "eager_step_ms_p99": float(np.percentile(eper, 99)),
        "status_ok": int(np.sum(st == 0)),
        "graph_vs_eager_max_abs_dlogL": float(np.max(np.abs(ll - want))),
    }



def bench_da2d(args):
    """The two-band delayed-acceptance leg at full size: the 2D likelihood
    of the 2D leg (2048^2 FP64 grid, 1000x1000 unbinned pixel pairs), a
    64-walker stretch move (32 proposals a half-step, the 2D batch size).
    A plain burn-in of 150 steps, then the (theta, exact logL) pairs of the
    next 240 stored steps train a 12 -> 3x256 -> 1 surrogate (band-1 knots,
    colour sigma and offset knots); from that state a plain chain and a
    delayed-acceptance chain of the same length run and are compared on
    effective samples per second and time per step."""
    from pofd_affine_amd import _lib, synth2d, workload2d
    n, W, S = args.n_2d, 64, args.da2d_steps
    b1, b2, pix = synth2d.gaussian_beams()
    d1, d2, nsrc = synth2d.make_maps(args.side, seed=99, beams=(b1, b2, pix))
    inp = dict(beam1=b1, beam2=b2, pixsize=pix, d1=d1, d2=d2,
               minbeamval=1e-6, nbeamhist=120, nsrc=nsrc, nbins=None)
    ctx = workload2d.build_context2d(inp, n, synth2d.true_theta())
    th0 = synth2d.walker_thetas(W, seed=4242, scatter=0.01)
    npar, nfit = th0.shape[1], 12
    state = np.zeros(npar, dtype=np.uint8)
    state[nfit:nfit + 2] = 1      # sigma multipliers fixed
    state[nfit + 2:] = 2          # bonus: the counts integrals
    fit = list(range(nfit))

    def make(pos, seed):
        e = _lib.Ensemble(ctx, _lib.TARGET_2D, W, npar, seed=seed)
        e.set_param_state(state)
        e.set_exponent_mode(True)
        e.set_positions(pos)
        return e

    ens = make(th0, 61)
    t0 = time.perf_counter()
    ens.run(150, thin=0)
    ens.run(240, thin=1)
    t_burn = time.perf_counter() - t0
    ch, cll = ens.chain()
    x = ch[:, :, :nfit].reshape(-1, nfit)
    y = cll.reshape(-1)
    start, _, _, _ = ens.state()
    del ens
    net = _lib.Mlp(nfit, 256, seed=15)
    t0 = time.perf_counter()
    tr, ho = net.train(x, y, epochs=150, batch=256, lr=1e-3,
                       holdout_frac=0.1, seed=16)
    t_train = time.perf_counter() - t0

    def chain(use_da):
        e = make(start, 71)
        on = e.set_surrogate(net, fit) if use_da else False
        e.run(3, thin=0)  # warm the launch path
        n0 = e.counts()
        e.clear_chain()
        t0 = time.perf_counter()
        e.run(S, thin=1)
        wall = time.perf_counter() - t0
        n1 = e.counts()
        taus = []
        for i in fit:
            try:
                taus.append(e.acor(i)[0])
            except _lib.PofdError:
                taus.append(float("nan"))
        tau = float(np.nanmean(taus)) if np.any(np.isfinite(taus)) else float("nan")
        return dict(on=bool(on), wall=wall, tau=tau,
                    nprop=int(n1[0] - n0[0]), nexact=int(n1[1] - n0[1]))

    plain = chain(False)
    da = chain(True)

    def side(r):
        ess = W * S / r["tau"] / r["wall"] if np.isfinite(r["tau"]) else float("nan")
        return {"wall_s": r["wall"], "ms_per_step": 1e3 * r["wall"] / S,
                "tau_mean": r["tau"], "ess_per_s": ess,
                "exact_evals": r["nexact"], "proposals": r["nprop"]}

    sp, sd = side(plain), side(da)
    return {
        "workload": "2D P(D) stretch move, %dx%d grid, %d walkers (32 "
                    "proposals a half-step), %d steps per chain, %dx%d "
                    "unbinned pixel pairs; surrogate MLP %d -> 3x256 GELU -> 1 "
                    "(bf16, fp32 accumulate) trained on %d (theta, logL) "
                    "pairs from 240 stored steps after 150 burn-in steps"
                    % (n, n, W, S, args.side, args.side, nfit, x.shape[0]),
        "surrogate": {"train_rms": tr, "holdout_rms": ho, "gate": 0.5,
                      "enabled": da["on"], "train_s": t_train},
        "burn_and_collect_s": t_burn,
        "plain": sp, "da": sd,
        "ess_per_s_ratio": sd["ess_per_s"] / sp["ess_per_s"],
        "step_time_ratio": plain["wall"] / da["wall"],
        "exact_fraction": da["nexact"] / max(da["nprop"], 1),
    }

def bench_da(args, inp):
    """The delayed-acceptance leg: the same binned 1D likelihood (nflux
    2^17), a 512-walker stretch move.  A plain burn-in of 500 steps from a
    0.01-dex ball around the generating knots, then the (theta,
    exact logL) pairs of the next 100 stored steps (51,200 pairs) train an
    8 -> 3x256 -> 1 surrogate on the GPU; from that one state a plain chain
    and a delayed-acceptance chain of the same length run, and are compared
    on effective samples per second (walkers x steps / tau of the
    walker-averaged series, per knot) and on the knot posterior medians; a
    second plain chain with another seed gives the Monte Carlo spread the
    median shift is to be read against.  A surrogate-transition chain
    (--st-steps surrogate stretch moves per walker and half-step, one exact
    correction at the end) runs from the same state and is compared the
    same way."""
    from pofd_affine_amd import _lib, synth, workload
    N, W = args.nflux, args.da_walkers
    ctx = workload.build_context(inp, N, args.ninterp, binned=True)
    th0 = synth.walker_thetas(W, seed=777, scatter=0.01)
    npar, nk = th0.shape[1], 8
    # workgroups per walker measured for the batch sizes a half-step meets
    # (the plain chain's 256 proposals, the delayed-acceptance survivors)
    tune_sizes = [16, 32, 48, 56, 64, 72, 80, 96, 128, 192, 256]
    t0 = time.perf_counter()
    tuned = ctx.autotune_parts(th0, tune_sizes, reps=5)
    t_tune = time.perf_counter() - t0
    state = np.zeros(npar, dtype=np.uint8)
    state[nk] = 1          # sigma multiplier fixed
    state[nk + 1:] = 2     # bonus: the counts integrals
    knots = list(range(nk))

    def make(pos, seed):
        e = _lib.Ensemble(ctx, _lib.TARGET_1D, W, npar, seed=seed)
        e.set_param_state(state)
        e.set_exponent_mode(True)
        e.set_positions(pos)
        return e

    ens = make(th0, 31)
    t0 = time.perf_counter()
    ens.run(args.da_burn, thin=0)
    ens.run(args.da_train_steps, thin=1)
    t_burn = time.perf_counter() - t0
    ch, cll = ens.chain()
    x = ch[:, :, :nk].reshape(-1, nk)
    y = cll.reshape(-1)
    start, _, _, _ = ens.state()
    del ens
    net = _lib.Mlp(nk, 256, seed=5)
    t0 = time.perf_counter()
    tr, ho = net.train(x, y, epochs=args.da_epochs, batch=512, lr=1e-3,
                       holdout_frac=0.1, seed=6)
    t_train = time.perf_counter() - t0

    def chain(use_da, seed=41, inner=1):
        e = make(start, seed)
        on = e.set_surrogate(net, knots) if use_da else False
        if on and inner > 1:
            e.set_surrogate_steps(inner)
        e.run(5, thin=0)  # warm the launch path
        n0 = e.counts()
        e.clear_chain()
        t0 = time.perf_counter()
        e.run(args.da_steps, thin=1)
        wall = time.perf_counter() - t0
        n1 = e.counts()
        c, _ = e.chain()
        taus = []
        for i in knots:
            try:
                taus.append(e.acor(i)[0])
            except _lib.PofdError:
                taus.append(float("nan"))
        _, _, nacc, nerr = e.state()
        return dict(on=on, wall=wall, chain=c[:, :, :nk],
                    taus=np.array(taus),
                    nprop=n1[0] - n0[0], nexact=n1[1] - n0[1],
                    acc=float(nacc.sum()) / (W * (args.da_steps + 5)))

    plain = chain(False)
    da = chain(True)
    control = chain(False, seed=43)  # the Monte Carlo spread of the medians
    st = chain(True, inner=args.st_steps) if args.st_steps > 1 else None
    S = args.da_steps

    def ess_per_s(r):
        tau = np.nanmean(r["taus"])
        return W * S / tau / r["wall"], tau

    ep, tp = ess_per_s(plain)
    ed, td = ess_per_s(da)
    flat_p = plain["chain"].reshape(-1, nk).astype(np.float64)
    flat_d = da["chain"].reshape(-1, nk).astype(np.float64)
    flat_c = control["chain"].reshape(-1, nk).astype(np.float64)
    sig = flat_p.std(axis=0)
    dmed = (np.median(flat_d, axis=0) - np.median(flat_p, axis=0)) / sig
    dctl = (np.median(flat_c, axis=0) - np.median(flat_p, axis=0)) / sig
    res = {
        "workload": "1D P(D) stretch move, nflux=%d, %d walkers, %d steps "
                    "per chain, binned %dx%d map; surrogate MLP %d -> "
                    "3x256 GELU -> 1 (bf16, fp32 accumulate) trained on "
                    "%d (theta, logL) pairs from %d stored steps after %d "
                    "burn-in steps" % (N, W, S, args.side, args.side, nk,
                                       x.shape[0], args.da_train_steps,
                                       args.da_burn),
        "surrogate": {"train_rms": tr, "holdout_rms": ho, "gate": 0.5,
                      "enabled": bool(da["on"]), "train_s": t_train,
                      "epochs": args.da_epochs},
        "autotune": {"sizes": tune_sizes,
                     "parts": [int(v) for v in tuned],
                     # the untuned rule: 256 // n clamped to 1..4, 3 -> 2
                     "size_rule": [{3: 2}.get(v, v) for v in
                                   (min(4, max(1, 256 // n))
                                    for n in tune_sizes)],
                     "seconds": t_tune},
        "burn_and_collect_s": t_burn,
        "plain": {"wall_s": plain["wall"], "ms_per_step":
                  1e3 * plain["wall"] / S, "tau_mean": tp,
                  "tau": plain["taus"].tolist(), "ess_per_s": ep,
                  "exact_evals": plain["nexact"],
                  "acceptance": plain["acc"]},
        "da": {"wall_s": da["wall"], "ms_per_step": 1e3 * da["wall"] / S,
               "tau_mean": td, "tau": da["taus"].tolist(), "ess_per_s": ed,
               "exact_evals": da["nexact"], "proposals": da["nprop"],
               "acceptance": da["acc"]},
        "ess_per_s_ratio": ed / ep,
        "step_time_ratio": plain["wall"] / da["wall"],
        # steps after which the one-off training (not the burn-in, which
        # both chains need) is paid back by the faster DA steps
        "breakeven_steps_incl_training": (
            t_train / ((plain["wall"] - da["wall"]) / S)
            if plain["wall"] > da["wall"] else None),
        "median_shift_sigma": dmed.tolist(),
        "max_abs_median_shift_sigma": float(np.max(np.abs(dmed))),
        # a second plain chain with another seed: how far two exact chains
        # of this length differ in their medians
        "control_median_shift_sigma": dctl.tolist(),
        "control_max_abs_median_shift_sigma": float(np.max(np.abs(dctl))),
        "control_tau_mean": float(np.nanmean(control["taus"])),
    }
    if st is not None:
        es, ts = ess_per_s(st)
        flat_s = st["chain"].reshape(-1, nk).astype(np.float64)
        dst = (np.median(flat_s, axis=0) - np.median(flat_p, axis=0)) / sig
        res["surrogate_transition"] = {
            "inner_steps": args.st_steps,
            "wall_s": st["wall"], "ms_per_step": 1e3 * st["wall"] / S,
            "tau_mean": ts, "tau": st["taus"].tolist(), "ess_per_s": es,
            "exact_evals": st["nexact"], "proposals": st["nprop"],
            "acceptance": st["acc"],
            "ess_per_s_ratio": es / ep,
            "median_shift_sigma": dst.tolist(),
            "max_abs_median_shift_sigma": float(np.max(np.abs(dst))),
        }
    return res


def bench_2d(args, hiprt):
    """The 2D leg: calcLikeDouble::getLogLike for a batch of walkers on a
    n x n grid (2048^2 FP64), 1000x1000 synthetic pixel pairs (unbinned),
    Gaussian beams of FWHM 18.1" and 25.2" at 6" in a 120x120-bin 2D
    histogram, sigma 2.5 / 2.7 mJy, 32 walkers per launch."""
    from pofd_affine_amd import synth2d, workload2d
    n, W = args.n_2d, args.walkers_2d
    b1, b2, pix = synth2d.gaussian_beams()
    d1, d2, nsrc = synth2d.make_maps(args.side, seed=99, beams=(b1, b2, pix))
    inp = dict(beam1=b1, beam2=b2, pixsize=pix, d1=d1, d2=d2,
               minbeamval=1e-6, nbeamhist=120, nsrc=nsrc, nbins=None)
    th = synth2d.walker_thetas(W, seed=4321)
    stride = th.shape[1]
    ctx = workload2d.build_context2d(inp, n, synth2d.true_theta())
    d_th = hiprt.DeviceBuffer(th.nbytes)
    d_th.upload(th)
    d_ll = hiprt.DeviceBuffer(W * 8)
    d_st = hiprt.DeviceBuffer(W * 4)
    stream = hiprt.Stream()
    for _ in range(2):
        ctx.loglike2d_batch_device(d_th.value, W, stride, d_ll.value,
                                   d_st.value, stream.ptr.value)
    stream.synchronize()
    steps = args.steps_2d
    evs = [hiprt.Event() for _ in range(steps + 1)]
    t0 = time.perf_counter()
    evs[0].record(stream)
    for i in range(steps):
        ctx.loglike2d_batch_device(d_th.value, W, stride, d_ll.value,
                                   d_st.value, stream.ptr.value)
        evs[i + 1].record(stream)
    stream.synchronize()
    wall = time.perf_counter() - t0
    per = [evs[i].elapsed_ms(evs[i + 1]) for i in range(steps)]
    # stage durations from two more batches with HIP events around every
    # stage (kept out of the timed batches)
    ctx.set_stage_timing(True)
    for _ in range(2):
        ctx.loglike2d_batch_device(d_th.value, W, stride, d_ll.value,
                                   d_st.value, stream.ptr.value)
    stream.synchronize()
    stage_ms, nl = ctx.stage_times()
    stage_ms = {k: v / max(nl, 1) for k, v in stage_ms.items()}
    ctx.set_stage_timing(False)
    ll = np.zeros(W)
    st = np.zeros(W, dtype=np.int32)
    d_ll.download(ll)
    d_st.download(st)
    batch_ms = float(np.mean(per))
    # algorithmic bytes of one evaluation: the three global stage
    # boundaries (R -> forward transform, spectrum -> inverse transform,
    # P(D) -> split/normalise/logL), each a write and a read of 8 n^2
    # bytes, plus the pixel pairs amortised over the batch
    b_eval = 48.0 * n * n + 16.0 * d1.size / W
    achieved = b_eval * W / (batch_ms * 1e-3) / 1e9
    res = {
        "value": W / (wall / steps), "unit": "evals/s",
        "workload": "2D P(D) log likelihood, %dx%d grid, %d walkers per "
                    "launch, %dx%d synthetic pixel pairs (unbinned), "
                    "120x120-bin 2D beam histogram, sigma 2.5/2.7 mJy"
                    % (n, n, W, args.side, args.side),
        "steps": steps, "batch_ms": batch_ms, "stage_ms": stage_ms,
        "status_ok": int(np.sum(st == 0)),
        "pipeline": {"bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "bytes_alg_per_eval": b_eval},
    }
    if not args.no_cpu and args.cpu_evals_2d > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle2d_side import build_oracle2d
        ne = min(args.cpu_evals_2d, W)
        nthr, model = host_cpu()
        nthr = args.cpu_threads or nthr
        L = build_oracle2d(inp, n, synth2d.true_theta())
        t0 = time.perf_counter()
        ref, rst = L.loglike_batch(th[:ne], nthr)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {
            "value": ne / dt, "unit": "evals/s", "cores": nthr,
            "kind": "port",
            "sample": "%d walkers of the same 2D workload on %d host "
                      "threads, one evaluation per thread at a time; own "
                      "radix-2 FFT (FFTW is absent)" % (ne, nthr),
            "cpu_model": model}
        ok = (rst == 0) & (st[:ne] == 0)
        res["parity_max_abs_dlogL"] = float(
            np.max(np.abs(ref[ok] - ll[:ne][ok]))) if ok.any() else None
    return res


if __name__ == "__main__":
    main()
