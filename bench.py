"""Benchmark: IFGF matvecs/s of the Helmholtz single-layer operator on one MI355X.

Workload (the headline configuration): N = 1,048,576 points uniformly random on the unit
sphere, targets = sources, kappa = 100, Chebyshev order 6, leaf size 64, complex N(0,1)
densities, at relative error <= 1e-4 against the direct sum: the IFGF tolerance 1e-5 lets the
order estimate (ifgfoperator.hpp:108-244) raise the orders to (6, 8, 7), which gives 6.4e-5;
order 6 without a tolerance gives 1.35e-3 and is timed beside it (`fixed_order`).  A step is
one mult (one matvec) with the densities already resident in HBM.  The mixed form (FP32
arithmetic, FP64 cone coordinates / phases / geometry) is the measured value; the plain FP64
form is timed beside it, and the relative error of both against an FP64 direct sum on
sampled targets is reported.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def sphere(n, seed):
    g = np.random.default_rng(seed).standard_normal((n, 3))
    return g / np.linalg.norm(g, axis=1)[:, None]


def direct_sum_torch(x, w, idx, kappa, dev):
    """FP64 direct sum sum_i w_i e^{i kappa d}/(4 pi d), d != 0, at targets x[idx] (torch)."""
    import torch
    X = torch.as_tensor(x, dtype=torch.float64, device=dev)
    W = torch.as_tensor(w, dtype=torch.complex128, device=dev)
    out = []
    for chunk in np.array_split(np.asarray(idx), max(1, len(idx) // 32)):
        Y = X[torch.as_tensor(chunk, device=dev)]
        diff = Y[:, None, :] - X[None, :, :]  # explicit differences: exact zero self-distance
        d = torch.sqrt((diff * diff).sum(dim=2))
        g = torch.where(d > 0, torch.exp(1j * kappa * d) / (4 * np.pi * torch.clamp(d, min=1e-300)),
                        torch.zeros((), dtype=torch.complex128, device=dev))
        out.append((g * W[None, :]).sum(dim=1))
    return torch.cat(out).cpu().numpy()


# the kernels of each phase, the first one found in a counter summary is the phase's
KERNEL_OF_PHASE = {"near": ("k_near",), "leaf": ("k_leaf_box", "k_leaf"),
                   "upward": ("k_up_grp", "k_up_cell"), "transform": ("k_transform_t",),
                   "far": ("k_far",)}
# what bounds each phase's kernel below the FP32 vector roofline, from the rocprofv3 SQ
# instruction and HBM counters of the 1M-point workloads (profiles/)
LIMITER = {
    "near": "VALU (issue rate near the practical ceiling)",
    "leaf": "VALU (issue rate near the practical ceiling)",
    "upward": "vector-memory latency of the dependent record / coefficient loads (about three "
              "quarters of the wave cycles wait), then VALU",
    "transform": "LDS traffic between the axis stages and the coefficient writes to HBM",
    "far": "vector-memory latency of the cell-table and coefficient gathers (more than half of "
           "the wave cycles wait) at 4 waves per SIMD, then VALU in the FP64 cone location",
}
# rocprofv3 counter summaries of the 1M-point workloads, by tolerance
COUNTERS = {
    1e-5: os.path.join(HERE, "profiles",
                       "counters_helmholtz_sphere_1M_kappa100_order6_tol1e-5_mixed.json"),
    -1.0: os.path.join(HERE, "profiles", "counters_helmholtz_sphere_1M_kappa100_order6_mixed.json"),
}


def load_counters(tol):
    """HBM bytes per mult and launches per mult of each kernel at a 1M-point workload, from
    the committed rocprofv3 counter summary (None when absent)."""
    try:
        with open(COUNTERS[tol]) as f:
            return json.load(f)
    except (KeyError, OSError, ValueError):
        return None


def flops_per_mult(work, transform_cones=None):
    """Algorithmic flops of one mult (transcendentals count 1 each).  transform_cones: the
    cones of the standalone transform kernel; when it is fewer than all, leaf sampling and the
    upward kernel transform their own cones, and those transform flops count for their phases
    (the leaf cones for leaf sampling, the rest for the upward pass)."""
    n_lo = work["n_lo"]
    cones = work["coarse_cones"]
    tc = cones if transform_cones is None else transform_cones
    tf = work["transform_flops_per_cone"]
    leaf_tf = work["leaf_cones"] if tc < cones else 0.0
    return {
        "near": 23.0 * work["near_pairs"],
        "leaf": 26.0 * work["leaf_pairs"] + tf * leaf_tf,
        "upward": (54.0 + 4.0 * n_lo) * work["up_evals"] + tf * (cones - tc - leaf_tf),
        "transform": tf * tc,
        "far": (44.0 + 4.0 * n_lo) * work["far_pairs"],
    }


def time_mults(op, w_dev, steps, warmup):
    import torch
    for _ in range(warmup):
        op.mult(w_dev)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        u = op.mult(w_dev)
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps, u


def host_cpu():
    """CPU model, logical CPUs of the host and CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, os.cpu_count() or 1, avail


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


# the CPU share of one GPU on the benchmark hosts: the baseline uses at most this many threads
CPU_SHARE = 16


def cpu_baseline(x, w, kappa, order, leaf, tol, stride, validate_n=0):
    """The FP64 CPU restatement of the reference (the oracle: the same plan, orders and
    recursive schedule, std::thread over the host's cores) on every stride-th subtree of the
    first level with more than 32 boxes (the reference's unit of parallel work).  Each sampled
    subtree runs on one thread; the sum of their durations divided by their fraction of the
    source points is the thread-seconds of a full mult, and the rate assumes a full mult keeps
    every thread busy (perfect scaling over the subtrees: generous to the CPU).  Two builds,
    sampled at the same time on half the threads each: -O3 for x86-64-v3 with -ffast-math
    (the reference's own flags; the reported value) and the same without fast-math (IEEE).
    validate_n > 0: also time full mults at validate_n points (kappa scaled by
    sqrt(validate_n / N), the same points per wavelength) against the same extrapolation."""
    import threading

    from oracle import pyoracle as po
    model, ncpu, avail = host_cpu()
    nthreads = max(1, min(CPU_SHARE, avail))
    half = max(1, nthreads // 2)
    orcs = {}
    t_init = {}
    for variant in ("_fast", "_o3"):
        orcs[variant] = po.Oracle(po.HELMHOLTZ, kappa, leaf, order, 1, tol, variant=variant)
        t0 = time.time()
        orcs[variant].init(x, x, nthreads=nthreads)
        t_init[variant] = time.time() - t0
        log(f"cpu baseline {variant}: plan {t_init[variant]:.1f}s")
    res = {}

    def run(variant):
        res[variant] = orcs[variant].mult_sample(w, stride, nthreads=half)

    t0 = time.time()
    ths = [threading.Thread(target=run, args=(v,)) for v in orcs]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    wall = time.time() - t0
    for v, (frac, ts) in res.items():
        log(f"cpu baseline {v}: {frac * 100:.1f}% of the points, {ts:.1f} thread-s -> "
            f"{ts / max(frac, 1e-12) / nthreads:.0f}s per mult on {nthreads} threads")
    del orcs
    frac, ts = res["_fast"]
    full_ts = ts / max(frac, 1e-12)
    out = {
        "value": nthreads / full_ts,
        "unit": "matvecs/s",
        "cores": nthreads,
        "kind": "port",
        "sample": (f"FP64 oracle -O3 -march=x86-64-v3 -ffast-math: the recursive phase of every "
                   f"{stride}th subtree at the first level with >32 boxes ({frac * 100:.1f}% of "
                   f"the source points, one thread per subtree) took {ts:.1f} thread-seconds; a "
                   f"full mult = that / the point fraction / {nthreads} threads (perfect "
                   f"scaling assumed); both builds sampled at once in {wall:.1f}s wall; plan "
                   f"build {t_init['_fast']:.1f}s untimed"),
        "ieee_o3_value": nthreads * res["_o3"][0] / max(res["_o3"][1], 1e-12),
        "cpu_model": model,
        "nproc": ncpu,
        "cpus_available": avail,
    }
    if validate_n > 0:
        n = int(validate_n)
        kv = kappa * np.sqrt(n / len(x))
        xv = sphere(n, 1099)
        wv = np.random.default_rng(2099).standard_normal(n) + \
            1j * np.random.default_rng(2199).standard_normal(n)
        orc = po.Oracle(po.HELMHOLTZ, kv, leaf, order, 1, tol, variant="_fast")
        orc.init(xv, xv, nthreads=nthreads)
        orc.mult(wv, nthreads=nthreads)
        ts_full = []
        for _ in range(3):
            t0 = time.time()
            orc.mult(wv, nthreads=nthreads)
            ts_full.append(time.time() - t0)
        fv, tv = orc.mult_sample(wv, stride, nthreads=nthreads)
        out["validation"] = {"n_points": n, "kappa": kv,
                             "full_mult_s_median": float(np.median(ts_full)),
                             "extrapolated_s": tv / max(fv, 1e-12) / nthreads,
                             "sample_fraction": fv}
        log(f"cpu baseline validation at {n} points: {out['validation']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--kappa", type=float, default=100.0)
    ap.add_argument("--order", type=int, default=6)
    ap.add_argument("--leaf", type=int, default=64)
    ap.add_argument("--tol", type=float, default=1e-5,
                    help="IFGF tolerance: > 0 lets the order estimate raise the orders "
                         "(ifgfoperator.hpp:108-244); <= 0 keeps the fixed order")
    ap.add_argument("--no-fixed-order", action="store_true",
                    help="skip the run of the same order without a tolerance")
    ap.add_argument("--samples", type=int, default=1000)
    ap.add_argument("--cpu-stride", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-validate", type=int, default=0, metavar="N",
                    help="also time full CPU mults at N points against the extrapolated sample")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes (no tuning after init)")
    ap.add_argument("--launch-shape", type=str, default=None,
                    help="JSON dict of launch shapes to apply instead of tuning, e.g. "
                         "'{\"xcd_run_log2\": 4, \"up_cones\": 16}'")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device once (a plain copy kernel of "
                         "known size, for scaling profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build "
              "shards nothing across GPUs)", file=sys.stderr)
        sys.exit(2)

    import torch
    import ifgf_amd

    dev = torch.device("cuda:0")
    if args.copy_calibration > 0:
        a = torch.empty(args.copy_calibration << 18, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    N = args.n
    x = sphere(N, 1002)
    rng = np.random.default_rng(2002)
    w = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    idx = np.random.default_rng(3002).choice(N, size=min(args.samples, N), replace=False)
    print(f"# N={N} kappa={args.kappa} order={args.order} leaf={args.leaf}", file=sys.stderr,
          flush=True)

    t0 = time.time()
    op = ifgf_amd.HelmholtzIfgfOperator(args.kappa, args.leaf, args.order, 1, args.tol,
                                        precision="mixed")
    op.init(x.T, x.T)
    t_init = time.time() - t0
    work = op.plan_work()
    # the cones of the standalone transform (fewer when the upward kernel transforms its own)
    transform_cones = op.plan_counts()["transform_cones"]
    orders = {k: v.tolist() for k, v in op.plan_orders().items()}
    print(f"# init {t_init:.1f}s  {op.plan_counts()}  orders {orders}", file=sys.stderr,
          flush=True)
    # launch shapes: timed on this operator once after init (untimed, like init)
    t0 = time.time()
    if args.launch_shape:
        op.launch_shape = json.loads(args.launch_shape)
        shape = op.launch_shape
    elif args.no_autotune:
        shape = op.launch_shape
    else:
        shape = op.autotune(nrhs=1, reps=3)
    t_tune = time.time() - t0
    print(f"# launch shape {shape} (autotune {t_tune:.1f}s)", file=sys.stderr, flush=True)
    w_dev = torch.as_tensor(w, dtype=torch.complex128, device=dev)
    ms, u = time_mults(op, w_dev, args.steps, args.warmup)
    print(f"# mixed: {ms:.3f} ms/matvec", file=sys.stderr, flush=True)

    # the same mults replayed as the operator's captured HIP graph (kernels between the gather
    # and the scatter)
    op.use_graph = True
    ms_graph, _ = time_mults(op, w_dev, args.steps, args.warmup)
    op.use_graph = False
    print(f"# mixed, graph replay: {ms_graph:.3f} ms/matvec", file=sys.stderr, flush=True)

    # per-phase device times: the timed steps repeated with the library's phase events (HIP
    # events on the mult's stream around every launch group, read once per mult)
    os.environ["IFGF_PHASE_TIMING"] = "1"
    acc = None
    for _ in range(args.steps):
        op.mult(w_dev)
        t = op.last_timings()
        acc = dict(t) if acc is None else {k: acc[k] + t[k] for k in acc}
    torch.cuda.synchronize()
    del os.environ["IFGF_PHASE_TIMING"]
    phases = {k: float(v) / args.steps for k, v in acc.items()}
    print(f"# phases (ms, mean of {args.steps}): {phases}", file=sys.stderr, flush=True)

    ref = direct_sum_torch(x, w, idx, args.kappa, dev)
    u_np = u.cpu().numpy()
    err = float(np.linalg.norm(u_np[idx] - ref) / np.linalg.norm(ref))

    fp64 = None
    if not args.no_fp64:
        op64 = ifgf_amd.HelmholtzIfgfOperator(args.kappa, args.leaf, args.order, 1, args.tol,
                                              precision="fp64")
        op64.init(x.T, x.T)
        if not (args.no_autotune or args.launch_shape):
            op64.autotune(nrhs=1, reps=2)
        ms64, u64 = time_mults(op64, w_dev, max(2, args.steps // 3), 1)
        u64_np = u64.cpu().numpy()
        fp64 = {
            "ms_per_step": ms64,
            "value": 1000.0 / ms64,
            "rel_err_vs_direct": float(np.linalg.norm(u64_np[idx] - ref) / np.linalg.norm(ref)),
            "mixed_vs_fp64_rel_diff": float(np.linalg.norm(u_np - u64_np) / np.linalg.norm(u64_np)),
        }
        print(f"# fp64: {ms64:.3f} ms/matvec", file=sys.stderr, flush=True)
        del op64
        torch.cuda.empty_cache()

    # the same order without a tolerance (orders (3, 6, 6) / (2, 3, 3): faster, less accurate)
    fixed = None
    if not args.no_fixed_order and args.tol > 0:
        del op
        torch.cuda.empty_cache()
        opf = ifgf_amd.HelmholtzIfgfOperator(args.kappa, args.leaf, args.order, 1, -1,
                                             precision="mixed")
        t0 = time.time()
        opf.init(x.T, x.T)
        tf_init = time.time() - t0
        shape_f = (opf.launch_shape if (args.no_autotune or args.launch_shape)
                   else opf.autotune(nrhs=1, reps=3))
        msf, uf = time_mults(opf, w_dev, args.steps, args.warmup)
        uf_np = uf.cpu().numpy()
        fixed = {
            "tolerance": -1.0,
            "orders": {k: v.tolist() for k, v in opf.plan_orders().items()},
            "ms_per_step": msf,
            "value": 1000.0 / msf,
            "rel_err_vs_direct": float(np.linalg.norm(uf_np[idx] - ref) / np.linalg.norm(ref)),
            "init_s": tf_init,
            "launch_shape": shape_f,
        }
        print(f"# fixed order: {msf:.3f} ms/matvec", file=sys.stderr, flush=True)
        del opf
        torch.cuda.empty_cache()

    fl = flops_per_mult(work, transform_cones)
    full_size = (N, args.kappa, args.order, args.leaf) == (1048576, 100.0, 6, 64)
    counters = load_counters(args.tol) if full_size else None
    names = {"near": "near", "leaf": "leaf", "upward": "upward", "transform": "transform",
             "far": "far"}
    dom = max(names, key=lambda k: phases.get(k, 0.0))
    t_dom = phases[dom] * 1e-3
    achieved = fl[dom] / t_dom / 1e12 if t_dom > 0 else 0.0
    traffic = None
    kc = None
    if counters:
        found = [k for k in KERNEL_OF_PHASE[dom] if k in counters.get("kernels", {})]
        kc = counters["kernels"][found[0]] if found else None
    if kc:
        traffic = (kc["fetch_bytes_per_mult"] + kc["write_bytes_per_mult"]) / kc["launches_per_mult"]
    roofline = {
        "bound": "valu",
        "limiter": LIMITER[dom],
        "achieved": achieved,
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP32_TFLOPS,
        "traffic": traffic,
        "kernel": dom,
        "launches_per_mult": kc["launches_per_mult"] if kc else None,
        "note": "no kernel issues MFMA; the roofline is the FP32 vector rate, peak = MI355X FP32 "
                "157.3 TFLOP/s (with packed FP32; the kernels are built without packed FP32 "
                "instructions, so their vector ceiling is 78.6): algorithmic flops of the "
                "phase per mult / its device time "
                "(= per launch / mean launch duration); traffic = HBM bytes per launch from "
                "rocprofv3 FETCH_SIZE/WRITE_SIZE scaled by a plain copy (" +
                (counters.get("source", "") if counters else "no counter profile") + ")",
        "phase_ms": phases,
        "phase_tflops": {k: (fl[k] / (phases[k] * 1e-3) / 1e12 if phases.get(k, 0) > 0 else None)
                         for k in fl},
    }

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(x, w, args.kappa, args.order, args.leaf, args.tol, args.cpu_stride,
                           args.cpu_validate)

    line = {
        "metric": "IFGF matvecs/sec at N=1M, kappa=100, rel.err<=1e-4 vs FP64 direct; "
                  "FP32-mixed vs plain FP64",
        "value": 1000.0 / ms,
        "unit": "matvecs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 cone coordinates, phases and geometry)",
        "data": "synthetic",
        "config": {
            "workload": "helmholtz_single_layer_sphere",
            "n_points": N,
            "kappa": args.kappa,
            "order": args.order,
            "leaf_size": args.leaf,
            "tolerance": args.tol,
            "orders": orders,
            "targets": "= sources",
            "densities": "complex N(0,1), seed 2002; points seed 1002",
            "parallelism": "single GPU",
        },
        "graph_ms_per_step": ms_graph,
        "rel_err_vs_direct": err,
        "rel_err_samples": int(len(idx)),
        "init_s": t_init,
        "launch_shape": shape,
        "autotune_s": t_tune,
        "fp64_form": fp64,
        "fixed_order": fixed,
        "mixed_over_fp64": (fp64["ms_per_step"] / ms) if fp64 else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
