"""Benchmark of the DQMC sweep path on one MI355X.

Workload: the O(3) spin-fermion model at L=10 (N=100, n=400), beta=20,
delta_tau=0.1 (M=200 slices), Assaad checkerboard, safe_mult=10, hoppings
(1.0,0.5,0.5,1.0), mu=-0.5, lambda=1, r=1, c=3, u=1, box 0.5, seeded uniform
phi in [-1,1]^3. One step = one sweep = M update steps (propagate + local
updates of every site), the reference's printed "sweep duration"
(dqmc_framework.jl:259-262). Prints one JSON line.

The JSON line also carries the other legs of the metric: "batch" (aggregate sweeps/s of 64
independent chains of the same workload advanced together, with the roofline of its dominant
kernel), "batch_mixed" (aggregate sweeps/s of the 64-chain mixed batch: L in {6,8,10,12},
beta in {5,10,20}, per-chain lambda and r) and "grad" (dS/dphi evaluations per second at
L=12, beta=10).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--batch C]
                       [--batch-sweeps S] [--batch-mixed C] [--grad-evals E] [--no-tune]

Both sweep legs run the autotuner (dqmc_autotune) after init, outside the timed region;
its choice is reported under "tuning".
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sweeps/sec at L=10 β=20 (1 vs 64 chains); ∂S/∂φ evals/sec; at ΔG_max<1e-8"
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6


def workload_params(amd, L=10, slices=200, seed=1001):
    return amd.Params(L=L, slices=slices, delta_tau=0.1, safe_mult=10, hoppings=(1.0, 0.5, 0.5, 1.0), mu1=-0.5,
                      mu2=-0.5, lam=1.0, r=1.0, c=3.0, u=1.0, checkerboard=True, box=0.5, seed=seed)


def algorithmic_work(cls, n, acc_per_launch):
    """(bound, amount per launch, unit) for a kernel class; flop/byte model of the path:
    fused B application 32 n^2 bytes; accepted rank-4 update 32 n^2 bytes (read+write G);
    complex QR with pivoting 16/3 n^3 flop, Q formation 16/3 n^3, GEMM 8 n^3, inverse 8 n^3."""
    from dqmc_amd._capi import K_APPLY_B, K_GEMM, K_INVERSE, K_LOCAL, K_QFORM, K_QRP
    if cls == K_LOCAL:
        return "hbm", 32.0 * n * n * acc_per_launch, "B"
    if cls == K_APPLY_B:
        return "hbm", 32.0 * n * n, "B"
    if cls in (K_QRP, K_QFORM):
        return "mfma", 16.0 / 3.0 * n ** 3, "flop"
    if cls in (K_GEMM, K_INVERSE):
        return "mfma", 8.0 * n ** 3, "flop"
    return "hbm", 32.0 * n * n, "B"


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_L10_beta20_single_chain.json")


def measured_traffic(kernel_class):
    """HBM bytes per launch of a kernel class at this workload, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (scaled by a copy kernel of known size); None if absent."""
    try:
        with open(TRAFFIC_FILE) as f:
            k = json.load(f)["kernels"].get(kernel_class)
        return None if k is None else float(k["hbm_bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p_amd, phi, steps=20, repeats=5):
    """The oracle (numpy + scipy's OpenBLAS LAPACK: zgeqp3, zungqr, zgetrf, zgetri) on the same
    workload after its init, at 1 BLAS thread (the reference's default, OMP_NUM_THREADS=1)
    and at one thread per host core: `repeats` samples of `steps` consecutive update steps
    (each sample holds steps / safe_mult stack recomputes, the sweep's proportion) after one
    warm-up sample; the median sample is scaled to sweeps/s. The faster setting is `value`."""
    import scipy.linalg  # noqa: F401  (loads the BLAS so threadpoolctl can see it)
    from threadpoolctl import threadpool_info, threadpool_limits
    from oracle import dqmc_oracle as O
    p = O.Params(p_amd.L, p_amd.slices, delta_tau=p_amd.delta_tau, safe_mult=p_amd.safe_mult,
                 hoppings=p_amd.hoppings, mu1=p_amd.mu1, mu2=p_amd.mu2, lam=p_amd.lam, r=p_amd.r, c=p_amd.c,
                 u=p_amd.u, chkr=True, box=p_amd.box)
    blas = [{"api": b.get("internal_api"), "version": b.get("version"), "arch": b.get("architecture")}
            for b in threadpool_info() if b.get("user_api") == "blas"]
    ncores = os.cpu_count() or 1
    res = {}
    for threads in sorted({1, ncores}):
        with threadpool_limits(threads):
            o = O.DQMC(p, phi)
            o.init()
            stream = O.UniformStream(O.host_uniforms(p_amd.seed + 7, 4 * p.N * steps * (repeats + 1) + 16))
            samples = []
            for r in range(repeats + 1):
                t0 = time.perf_counter()
                for _ in range(steps):
                    o.update(stream)
                if r > 0:
                    samples.append(time.perf_counter() - t0)
        med = float(np.median(samples))
        res[threads] = {"value": (steps / p.slices) / med, "cores": threads,
                        "samples_s": [round(x, 4) for x in samples]}
    best = max(res.values(), key=lambda r: r["value"])
    return {"value": best["value"], "unit": "sweeps/s", "cores": best["cores"], "kind": "port",
            "sample": "median of %d samples of %d consecutive update steps (%d stack recomputes each) of the "
                      "L=10 beta=20 chain after init and one warm-up sample, numpy/OpenBLAS oracle, scaled to "
                      "sweeps/s (1 sweep = %d steps)" % (repeats, steps, steps // p.safe_mult, p.slices),
            "by_threads": {str(k): v for k, v in res.items()}, "cpu_model": _cpu_model(), "nproc": ncores,
            "blas": blas}


def mixed_params(amd, nchains=64):
    """The mixed batch: chain c has L = [6,8,10,12][c mod 4], beta = [5,10,20][c//4 mod 3],
    r = [-1,0,1][c//12 mod 3], lambda = 0.5 + 1.5 u_c (u_c from seed 2000 + c)."""
    ps = []
    for c in range(nchains):
        L = (6, 8, 10, 12)[c % 4]
        beta = (5, 10, 20)[(c // 4) % 3]
        p = workload_params(amd, L=L, slices=int(round(beta / 0.1)), seed=2000 + c)
        p.r = (-1.0, 0.0, 1.0)[(c // 12) % 3]
        p.lam = 0.5 + 1.5 * float(np.random.default_rng(2000 + c).uniform())
        ps.append(p)
    return ps


def batch_measure(amd, params, sweeps, warmup, tune=True, roofline=True):
    """Aggregate sweeps/s of independent chains advanced together by one batch (chains of one
    shape share every kernel launch; each shape group's sweep is replayed from HIP graphs).
    `warmup` >= 2 untimed sweeps capture both graphs (up and down sweeps) before timing. With
    `roofline` (one shape group), one more untimed eager sweep times the local-update kernel
    with HIP events for the roofline of the batch's dominant kernel."""
    from dqmc_amd import Batch
    from dqmc_amd import _capi
    nchains = len(params)
    b = Batch(params)
    b.use_graph(True)
    confs = [np.random.default_rng(p.seed).uniform(-1.0, 1.0, (3, p.N, p.slices)) for p in params]
    b.init(confs)
    tuning = b.autotune(reps=2) if tune else None
    if warmup > 0:
        b.sweep(warmup)
    b.synchronize()
    t0 = time.perf_counter()
    b.sweep(sweeps, sync=False)
    b.synchronize()
    dt = time.perf_counter() - t0
    sts = [b.stats(c) for c in range(nchains)]
    out = {"chains": nchains, "value": nchains * sweeps / dt, "unit": "sweeps/s (all chains)",
           "per_chain": sweeps / dt, "sweeps": sweeps, "warmup": warmup, "seconds": dt,
           "drift_max": max(st.drift_max for st in sts),
           "acceptance": sum(st.accepted for st in sts) / max(sum(st.proposed for st in sts), 1),
           "shapes": sorted({(p.L, p.slices) for p in params}), "tuning": tuning}
    if roofline and len(out["shapes"]) == 1:
        n = params[0].n
        acc0 = sum(st.accepted for st in sts)
        b.lib.dqmc_batch_profile(b.h, 1 << _capi.K_LOCAL)
        b.sweep(1)
        ms = np.zeros(_capi.K_NCLASS)
        cnt = np.zeros(_capi.K_NCLASS, dtype=np.int64)
        _capi.check(b.lib.dqmc_batch_profile_read(b.h, _capi.dptr(ms), cnt.ctypes.data_as(_capi._LL),
                                                  _capi.K_NCLASS))
        b.lib.dqmc_batch_profile(b.h, 0)
        acc = sum(b.stats(c).accepted for c in range(nchains)) - acc0
        launches = int(cnt[_capi.K_LOCAL])
        avg_ms = float(ms[_capi.K_LOCAL]) / max(launches, 1)
        work = 32.0 * n * n * acc / max(launches, 1)
        achieved = work / (avg_ms * 1e-3) / 1e9
        # 64 chains per launch move G through the windowed updates far below the unblocked
        # update's 32 n^2 bytes per acceptance (that rate exceeds the HBM peak here), so the
        # batch leg is priced by the update's arithmetic instead: 34 n^2 flop per acceptance
        flops = 34.0 * n * n * acc / max(launches, 1)
        tflops = flops / (avg_ms * 1e-3) / 1e12
        out["roofline"] = {"bound": "mfma", "kernel": "local_updates (all chains per launch)", "achieved": tflops,
                           "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": tflops / PEAK_FP64_TFLOPS,
                           "work_per_launch": flops, "work_unit": "flop", "avg_launch_ms": avg_ms,
                           "launches": launches, "unblocked_update_bytes_rate_GBs": achieved}
    b.close()
    return out


def grad_measure(amd, evals, warmup):
    """dS/dphi evaluations per second at L=12, beta=10 (n=576, M=100): one stabilised pass over
    all slices with the trace formula at every (component, site, slice).

    The warm-up evaluations run with every kernel class timed by HIP events on the chain's
    stream; they give the dominant class, its average launch time (the roofline entry, with
    the flop/byte model of algorithmic_work) and the evaluation's algorithmic FP64 work: the
    flop-bound classes' launches per evaluation times their model work. The timed
    evaluations run without events."""
    from dqmc_amd import DQMC, UniformStream, _capi
    p = workload_params(amd, L=12, slices=100, seed=3001)
    phi = np.random.default_rng(p.seed).uniform(-1.0, 1.0, (3, p.N, p.slices))
    c = DQMC(p, uniforms=UniformStream(p.seed))
    c.init(phi)
    warmup = max(warmup, 1)
    ms = np.zeros(_capi.K_NCLASS)
    cnt = np.zeros(_capi.K_NCLASS, dtype=np.int64)
    _capi.check(c.lib.dqmc_profile_read(c.h, None, None, 0))  # drop the init's records
    c.lib.dqmc_profile(c.h, (1 << _capi.K_NCLASS) - 1)
    for _ in range(warmup):
        c.grad()
    _capi.check(c.lib.dqmc_profile_read(c.h, _capi.dptr(ms), cnt.ctypes.data_as(_capi._LL), _capi.K_NCLASS))
    c.lib.dqmc_profile(c.h, 0)
    c.synchronize()
    t0 = time.perf_counter()
    for _ in range(evals):
        g, _s = c.grad()
    dt = time.perf_counter() - t0
    c.close()
    dom = int(np.argmax(ms))
    avg_ms = float(ms[dom]) / max(int(cnt[dom]), 1)
    bound, work, unit = algorithmic_work(dom, p.n, 0.0)
    if bound == "hbm":
        achieved, peak, runit = work / (avg_ms * 1e-3) / 1e9, PEAK_HBM_GBS, "GB/s"
    else:
        achieved, peak, runit = work / (avg_ms * 1e-3) / 1e12, PEAK_FP64_TFLOPS, "TFLOP/s"
    flop = 0.0
    for k in range(_capi.K_NCLASS):
        b, w, _u = algorithmic_work(k, p.n, 0.0)
        if b == "mfma" and cnt[k]:
            flop += w * float(cnt[k]) / warmup
    eval_tflops = flop * evals / dt / 1e12
    return {"value": evals / dt, "unit": "evals/s", "config": "L=12 beta=10 dtau=0.1 (M=100), n=576",
            "components": int(g.size), "evals": evals, "seconds": dt,
            "roofline": {"bound": bound, "kernel": _capi.K_NAMES[dom], "achieved": achieved, "peak": peak,
                         "unit": runit, "frac": achieved / peak, "work_per_launch": work, "work_unit": unit,
                         "avg_launch_ms": avg_ms, "launches_per_eval": int(cnt[dom]) // warmup},
            "eval_flop": {"flop_per_eval": flop, "achieved": eval_tflops, "peak": PEAK_FP64_TFLOPS,
                          "unit": "TFLOP/s", "frac": eval_tflops / PEAK_FP64_TFLOPS},
            "kernel_ms_per_eval_warmup": {_capi.K_NAMES[k]: round(float(ms[k]) / warmup, 3)
                                          for k in range(_capi.K_NCLASS)}}


def calibration_copy(mib):
    """One device-to-device copy of `mib` MiB (the runtime's copy kernel): a kernel of known
    HBM traffic that calibrates FETCH_SIZE / WRITE_SIZE counter passes. Uses the HIP runtime
    the library links (already loaded), not a second runtime instance."""
    import ctypes
    from dqmc_amd import _capi
    _capi.load()
    hip = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    nbytes = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(src), nbytes) or hip.hipMalloc(ctypes.byref(dst), nbytes):
        raise RuntimeError("calibration copy: hipMalloc failed")
    hip.hipMemset(src, 1, nbytes)
    if hip.hipMemcpy(dst, src, nbytes, 3) != 0:  # hipMemcpyDeviceToDevice
        raise RuntimeError("calibration copy: hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20, help="update steps per CPU-baseline sample")
    ap.add_argument("--batch", type=int, default=64, help="chains of the batch leg (0 skips it)")
    ap.add_argument("--batch-sweeps", type=int, default=5)
    ap.add_argument("--batch-warmup", type=int, default=2)
    ap.add_argument("--batch-mixed", type=int, default=64, help="chains of the mixed-batch leg (0 skips it)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (kernel profiles use this so tuning launches stay out)")
    ap.add_argument("--grad-evals", type=int, default=10, help="0 skips the gradient leg")
    ap.add_argument("--skip-single", action="store_true",
                    help="skip the single-chain leg (kernel profiles of the batch legs); value is then null")
    ap.add_argument("--pmc-copy-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device (a kernel of known traffic that "
                         "calibrates FETCH_SIZE / WRITE_SIZE counter runs); 0 skips it")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: a DQMC chain does not shard across GPUs and independent "
              "chains need no communication; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import dqmc_amd
    from dqmc_amd import DQMC, UniformStream
    from dqmc_amd import _capi

    if args.pmc_copy_mib > 0:
        calibration_copy(args.pmc_copy_mib)

    p = workload_params(dqmc_amd)
    rng = np.random.default_rng(p.seed)
    phi = rng.uniform(-1.0, 1.0, (3, p.N, p.slices))
    if args.skip_single:
        out = {"metric": METRIC, "value": None, "unit": "sweeps/s", "n_gpus": 1, "steps": 0,
               "warmup": 0, "higher_is_better": True, "note": "single-chain leg skipped"}
    else:
        chain = DQMC(p, uniforms=UniformStream(p.seed + 7))
        chain.init(phi)
        tuning = None if args.no_tune else chain.autotune(reps=2)
        # warmup with every kernel class timed, to find the dominant kernel
        chain.lib.dqmc_profile(chain.h, (1 << _capi.K_NCLASS) - 1)
        if args.warmup > 0:
            chain.sweep(args.warmup)
        ms = np.zeros(_capi.K_NCLASS)
        cnt = np.zeros(_capi.K_NCLASS, dtype=np.int64)
        _capi.check(chain.lib.dqmc_profile_read(chain.h, _capi.dptr(ms), cnt.ctypes.data_as(_capi._LL),
                                                _capi.K_NCLASS))
        dominant = int(np.argmax(ms)) if args.warmup > 0 else _capi.K_LOCAL
        warm_breakdown = {(_capi.K_NAMES[k]): round(float(ms[k]) / max(args.warmup, 1), 3) for k in range(len(ms))}
        chain.reset_stats()

        # timed region: only the dominant kernel class carries events
        chain.lib.dqmc_profile(chain.h, 1 << dominant)
        chain.synchronize()
        t0 = time.perf_counter()
        chain.sweep(args.steps, sync=False)
        chain.synchronize()
        dt = time.perf_counter() - t0
        st = chain.stats()
        _capi.check(chain.lib.dqmc_profile_read(chain.h, _capi.dptr(ms), cnt.ctypes.data_as(_capi._LL),
                                                _capi.K_NCLASS))
        avg_ms = float(ms[dominant]) / max(int(cnt[dominant]), 1)
        acc_per_slice = st.accepted / max(args.steps * p.slices, 1)
        bound, work, unit = algorithmic_work(dominant, p.n, acc_per_slice)
        if bound == "hbm":
            achieved = work / (avg_ms * 1e-3) / 1e9
            peak, runit = PEAK_HBM_GBS, "GB/s"
        else:
            achieved = work / (avg_ms * 1e-3) / 1e12
            peak, runit = PEAK_FP64_TFLOPS, "TFLOP/s"

        out = {
            "metric": METRIC,
            "value": args.steps / dt,
            "unit": "sweeps/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": dt / args.steps * 1e3,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f64",
            "data": "synthetic (seeded uniform phi in [-1,1]^3, host MT19937 uniform stream)",
            "config": {"workload": "O(3) spin-fermion DQMC sweep, single chain, L=10 beta=20 dtau=0.1 (M=200), "
                                   "Assaad checkerboard, safe_mult=10", "L": p.L, "n": p.n, "slices": p.slices,
                       "chains": 1},
            "roofline": {"bound": bound, "kernel": _capi.K_NAMES[dominant], "achieved": achieved, "peak": peak,
                         "unit": runit, "frac": achieved / peak,
                         "traffic": measured_traffic(_capi.K_NAMES[dominant]),
                         "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT),
                         "work_per_launch": work, "work_unit": unit, "avg_launch_ms": avg_ms,
                         "launches": int(cnt[dominant])},
            # the same launches priced by arithmetic (34 n^2 flop per accepted update)
            "roofline_flop": ({"achieved": 34.0 * p.n * p.n * acc_per_slice / (avg_ms * 1e-3) / 1e12,
                               "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                               "frac": 34.0 * p.n * p.n * acc_per_slice / (avg_ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS}
                              if dominant == _capi.K_LOCAL else None),
            "drift_max": st.drift_max,
            "acceptance": st.accepted / max(st.proposed, 1),
            "kernel_ms_per_sweep_warmup": warm_breakdown,
            "tuning": tuning,
        }
        chain.close()
    if args.batch > 0:
        params = [workload_params(dqmc_amd, seed=2000 + c) for c in range(args.batch)]
        out["batch"] = batch_measure(dqmc_amd, params, args.batch_sweeps, args.batch_warmup, tune=not args.no_tune)
        if out["value"]:
            out["batch"]["vs_single_chain"] = out["batch"]["value"] / out["value"]
    if args.batch_mixed > 0:
        # merged sweeps choose each launch's workgroups themselves: no autotuning
        out["batch_mixed"] = batch_measure(dqmc_amd, mixed_params(dqmc_amd, args.batch_mixed), args.batch_sweeps,
                                           args.batch_warmup, tune=False, roofline=False)
        if out["value"]:
            out["batch_mixed"]["vs_single_chain"] = out["batch_mixed"]["value"] / out["value"]
    if args.grad_evals > 0:
        out["grad"] = grad_measure(dqmc_amd, args.grad_evals, 1)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, phi, args.cpu_steps)
        if args.batch > 0:
            out["batch"]["vs_cpu"] = out["batch"]["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
