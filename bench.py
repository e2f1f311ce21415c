"""Benchmark of the DQMC sweep path on one MI355X.

Workload: the O(3) spin-fermion model at L=10 (N=100, n=400), beta=20,
delta_tau=0.1 (M=200 slices), Assaad checkerboard, safe_mult=10, hoppings
(1.0,0.5,0.5,1.0), mu=-0.5, lambda=1, r=1, c=3, u=1, box 0.5, seeded uniform
phi in [-1,1]^3. One step = one sweep = M update steps (propagate + local
updates of every site), the reference's printed "sweep duration"
(dqmc_framework.jl:259-262). Prints one JSON line.

The JSON line also carries the two other legs of the metric: "batch" (aggregate sweeps/s
of 64 independent chains of the same workload advanced together) and "grad" (dS/dphi
evaluations per second at L=12, beta=10).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--batch C]
                       [--batch-sweeps S] [--grad-evals E] [--no-tune]

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


def cpu_baseline(p_amd, phi, slices=100):
    """The oracle (numpy + OpenBLAS LAPACK, 1 BLAS thread) on `slices` update steps of the
    same workload after its init; scaled to sweeps/s."""
    from threadpoolctl import threadpool_limits
    from oracle import dqmc_oracle as O
    p = O.Params(p_amd.L, p_amd.slices, delta_tau=p_amd.delta_tau, safe_mult=p_amd.safe_mult,
                 hoppings=p_amd.hoppings, mu1=p_amd.mu1, mu2=p_amd.mu2, lam=p_amd.lam, r=p_amd.r, c=p_amd.c,
                 u=p_amd.u, chkr=True, box=p_amd.box)
    with threadpool_limits(1):
        o = O.DQMC(p, phi)
        o.init()
        stream = O.UniformStream(O.host_uniforms(p_amd.seed + 7, 4 * p.N * slices + 16))
        t0 = time.perf_counter()
        for _ in range(slices):
            o.update(stream)
        dt = time.perf_counter() - t0
    return {"value": (slices / p.slices) / dt, "unit": "sweeps/s", "cores": 1, "kind": "port",
            "sample": "%d of the %d update steps of one sweep (incl. %d stack recomputes) after init, "
                      "numpy/OpenBLAS oracle, 1 BLAS thread, scaled to sweeps/s" % (slices, p.slices,
                                                                                   slices // p.safe_mult)}


def batch_measure(amd, nchains, sweeps, warmup, tune=True):
    """Aggregate sweeps/s of `nchains` independent chains at the same workload (L=10, beta=20),
    advanced together by one batch (every kernel launch covers all chains; sweeps replayed
    from HIP graphs)."""
    from dqmc_amd import Batch
    params = [workload_params(amd, seed=2000 + c) for c in range(nchains)]
    b = Batch(params)
    b.use_graph(True)  # each chain group's sweep is replayed from a captured HIP graph
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
    b.close()
    return {"chains": nchains, "value": nchains * sweeps / dt, "unit": "sweeps/s (all chains)",
            "per_chain": sweeps / dt, "sweeps": sweeps, "warmup": warmup, "seconds": dt,
            "drift_max": max(st.drift_max for st in sts),
            "acceptance": sum(st.accepted for st in sts) / max(sum(st.proposed for st in sts), 1),
            "tuning": tuning}


def grad_measure(amd, evals, warmup):
    """dS/dphi evaluations per second at L=12, beta=10 (n=576, M=100): one stabilised pass over
    all slices with the trace formula at every (component, site, slice)."""
    from dqmc_amd import DQMC, UniformStream
    p = workload_params(amd, L=12, slices=100, seed=3001)
    phi = np.random.default_rng(p.seed).uniform(-1.0, 1.0, (3, p.N, p.slices))
    c = DQMC(p, uniforms=UniformStream(p.seed))
    c.init(phi)
    for _ in range(warmup):
        c.grad()
    c.synchronize()
    t0 = time.perf_counter()
    for _ in range(evals):
        g, _s = c.grad()
    dt = time.perf_counter() - t0
    c.close()
    return {"value": evals / dt, "unit": "evals/s", "config": "L=12 beta=10 dtau=0.1 (M=100), n=576",
            "components": int(g.size), "evals": evals, "seconds": dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-slices", type=int, default=100)
    ap.add_argument("--batch", type=int, default=64, help="chains of the batch leg (0 skips it)")
    ap.add_argument("--batch-sweeps", type=int, default=1)
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (kernel profiles use this so tuning launches stay out)")
    ap.add_argument("--grad-evals", type=int, default=3, help="0 skips the gradient leg")
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
        import torch
        src = torch.ones(args.pmc_copy_mib << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    p = workload_params(dqmc_amd)
    rng = np.random.default_rng(p.seed)
    phi = rng.uniform(-1.0, 1.0, (3, p.N, p.slices))
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
    chain.stream.advance(st.uniforms_used)
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
        "drift_max": st.drift_max,
        "acceptance": st.accepted / max(st.proposed, 1),
        "kernel_ms_per_sweep_warmup": warm_breakdown,
        "tuning": tuning,
    }
    chain.close()
    if args.batch > 0:
        out["batch"] = batch_measure(dqmc_amd, args.batch, args.batch_sweeps, 1, tune=not args.no_tune)
    if args.grad_evals > 0:
        out["grad"] = grad_measure(dqmc_amd, args.grad_evals, 1)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, phi, args.cpu_slices)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
