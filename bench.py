"""Headline benchmark: wall-seconds per topology-optimisation design iteration
of the top3D125 cantilever at 192x96x96 hex (1,769,472 elements, 5,447,811
DOFs, 5 multigrid levels, CG tolerance 1e-8) on one MI355X, with the
design loop (filter, SIMP, MGCG solve, sensitivities, OC) resident in HBM.

A step is one design iteration.  Prints one JSON line (rank 0, N = 1).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "gpu-to_amd"), os.path.join(ROOT, "oracle")]

METRIC = "wall-s/TO iter @1.77M hex, mixed+learned vs FP64 MGCG; matvec HBM GB/s"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
PMC_SUMMARY = os.path.join(ROOT, "profiles", "level0_operator_cantilever_192x96x96_pmc.md")
# variant -> (precision, level-0 kernel of the inner solve); kernel ids as
# gputo_params.l0_kernel (0 = FP64 node-gather for the plain path)
# name: (precision, level-0 kernel, sweeps before / after the coarse
# correction on levels 0 and 1 of the FP32 V-cycle (0 = the preset's 3 / 3),
# coarse_lag: V-cycles that use the previous design's coarsest factor[,
# level-1 sweeps before / after when they differ from level 0's[, sweeps
# of the levels from 2 on]])
VARIANTS = {
    "f64": (0, 0, 0, 0, 0),
    "mixed_gather": (1, 1, 0, 0, 0),  # FP32 inner solve, FP32 node-gather level-0 operator
    "mixed_mfma9": (2, 2, 0, 0, 0),   # MFMA level-0 operator, three-way BF16 split, all nine products
    "mixed_mfma": (2, 4, 0, 0, 0),    # MFMA level-0 operator, three-way BF16 split, six products
    "mixed_mfma3": (2, 3, 0, 0, 0),   # MFMA level-0 operator, two-way BF16 split, three products
    "mixed_mfma1": (2, 5, 0, 0, 0),   # MFMA level-0 operator, one BF16 product
    "mixed_mfma3s": (2, 3, 1, 3, 0),  # as mixed_mfma3 with 1 + 3 sweeps on levels 0 and 1
    "mixed_mfma3s3": (2, 3, 1, 3, 3),  # and the previous coarsest factor in the first 3 V-cycles
    "mixed_mfma3s11": (2, 3, 1, 1, 0),  # 1 + 1 sweeps on levels 0 and 1
    "mixed_mfma3s22": (2, 3, 2, 2, 0),  # 2 + 2 sweeps on levels 0 and 1
    "mixed_mfma3s11l3": (2, 3, 1, 1, 3),  # 1 + 1, previous coarsest factor in the first 3 V-cycles
    "mixed_mfma3s1113": (2, 3, 1, 1, 0, 1, 3),  # 1 + 1 on level 0, 1 + 3 on level 1
    "mixed_mfma3s1122": (2, 3, 1, 1, 0, 2, 2),  # 1 + 1 on level 0, 2 + 2 on level 1
    "mixed_mfma3s111": (2, 3, 1, 1, 0, 1, 1, 1, 1),  # 1 + 1 on every level
    "mixed_mfma3s1112": (2, 3, 1, 1, 0, 1, 1, 2, 2),  # 1 + 1 on levels 0 and 1, 2 + 2 below
    "mixed_mfma3s1114": (2, 3, 1, 1, 0, 1, 1, 4, 4),  # 1 + 1 on levels 0 and 1, 4 + 4 below
    "mixed_mfma3s1116": (2, 3, 1, 1, 0, 1, 1, 6, 6),  # 1 + 1 on levels 0 and 1, 6 + 6 below
    "mixed_mfma3s1118": (2, 3, 1, 1, 0, 1, 1, 8, 8),  # 1 + 1 on levels 0 and 1, 8 + 8 below
}


def level2_sweeps(v):
    """(pre, post) sweeps of the levels from 2 on (0, 0: the preset's)."""
    return (v[7], v[8]) if len(v) > 7 else (0, 0)


def level1_sweeps(v):
    """(pre, post) sweeps of level 1 of a variant tuple: its own entries when
    it has them, else level 0's."""
    return (v[5], v[6]) if len(v) > 5 else (v[2], v[3])


def pmc_traffic(kernel):
    """HBM bytes per launch of level-0 kernel `kernel` at 192x96x96 from the
    committed counter summary: FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the
    factors that make the same counters read 2 x 256 MiB for two 256 MiB
    device copies."""
    try:
        with open(PMC_SUMMARY) as f:
            for line in f:
                w = line.split()
                if len(w) == 3 and w[0] == "traffic_bytes" and w[1] == "kernel%d" % kernel:
                    return int(w[2])
    except OSError:
        pass
    return None


def level0_sweeps(name):
    """'pre+post' Jacobi sweeps on levels 0 and 1 of a mixed variant's FP32
    V-cycle (None for the FP64 variant, whose V-cycle is the preset's)."""
    v = VARIANTS[name.replace("+learned", "")]
    if v[0] == 0:
        return None
    n1 = level1_sweeps(v)
    s0 = "%d+%d" % (v[2] or 3, v[3] or 3)
    s1 = "%d+%d" % (n1[0] or 3, n1[1] or 3)
    out = s0 if s0 == s1 else "%s on level 0, %s on level 1" % (s0, s1)
    n2 = level2_sweeps(v)
    if n2[0] or n2[1]:
        out += ", %d+%d on the levels from 2" % (n2[0] or 3, n2[1] or 3)
    return out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_variant(gputo, name, args, Ke, F, fixed, x0):
    v = VARIANTS[name.replace("+learned", "")]
    prec, kernel, npre0, npost0, lag = v[:5]
    npre1, npost1 = level1_sweeps(v)
    npre2, npost2 = level2_sweeps(v)
    learned = name.endswith("+learned")
    s = gputo.Solver(args.nelx, args.nely, args.nelz, args.nl, precision=prec, learned=learned,
                     l0_kernel=kernel, mixed_npre0=npre0, mixed_npost0=npost0, mixed_npre1=npre1,
                     mixed_npost1=npost1, mixed_npre2=npre2, mixed_npost2=npost2, coarse_lag=lag)
    run = gputo.TopOpt(s, Ke, x0, F, fixed, volfrac=0.12, penal=3.0, rmin=np.sqrt(3.0), ft=1, move=0.2,
                       cg_tol=args.cg_tol, cg_maxit=args.cg_maxit)
    train_s = 0.0
    for w in range(args.warmup):
        st = run.iterate()
        log("[%s] warmup %d: C=%.6e nit=%d solve=%.3fs total=%.3fs" % (name, w, st["compliance"], st["cg_iters"],
                                                                       st["seconds_solve"], st["seconds_total"]))
        if learned:
            # train on the level-2 defects this solve logged (outside the timed steps)
            t0 = time.perf_counter()
            l0, l1 = s.learn_train(args.learn_steps, args.learn_lr)
            s.lib.gputo_device_synchronize()
            train_s += time.perf_counter() - t0
            log("[%s] training %d steps: energy %.4g -> %.4g (%.2fs)" % (name, args.learn_steps, l0, l1,
                                                                        time.perf_counter() - t0))
    s.lib.gputo_device_synchronize()
    t0 = time.perf_counter()
    hist = [run.iterate() for _ in range(args.steps)]
    s.lib.gputo_device_synchronize()
    wall = time.perf_counter() - t0
    for k, st in enumerate(hist):
        log("[%s] step %d: C=%.8e nit=%d relres=%.2e solve=%.3fs total=%.3fs" % (
            name, k, st["compliance"], st["cg_iters"], st["relres"], st["seconds_solve"], st["seconds_total"]))
    x, xp, U = run.state()
    out = {
        "s_per_iter": wall / args.steps,
        "solve_s_per_iter": float(np.mean([h["seconds_solve"] for h in hist])),
        "cg_iters": [h["cg_iters"] for h in hist],
        "compliance": [h["compliance"] for h in hist],
        "relres": [h["relres"] for h in hist],
        "learned_accepted": [h["learned_accepted"] for h in hist],
        "learned_rejected": [h["learned_rejected"] for h in hist],
        "train_s": train_s,
        "U": U,
    }
    # dominant kernel: the level-0 operator, timed with HIP events on the
    # solver's stream on the hierarchy of the last timed solve
    avg = ctypes.c_double()
    rc = s.lib.gputo_bench_operator(s.h, kernel, args.op_reps, ctypes.byref(avg))
    out["op_ms"] = avg.value if rc == 0 else None
    run.close()
    s.close()
    return out


def cpu_baseline(args, Ke, F, fixed, x0, nit_avg):
    """Oracle (C/OpenMP restatement of mgcg1.c + top3D125mod.m) on the host:
    two design iterations from x0 with the CG capped at 1 and 3 iterations
    give the fixed cost and the per-CG-iteration cost, extrapolated to the
    GPU's measured CG count (same algorithm, same iteration counts)."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "16"))
    oracle.set_num_threads(threads)
    times = {}
    for cap in (1, 3):
        st = oracle.DesignState(args.nelx, args.nely, args.nelz, x0)
        t0 = time.perf_counter()
        o = oracle.design_iter(st, Ke, F, fixed, args.nl, args.cg_tol, cap, 0.12, 3.0, np.sqrt(3.0), 1, 2.0, 0.2)
        times[cap] = time.perf_counter() - t0
        log("[cpu] design iteration with CG capped at %d: %.2fs (rc %d)" % (cap, times[cap], o["rc"]))
    t_it = (times[3] - times[1]) / 2
    t_fixed = times[1] - t_it
    return {
        "value": t_fixed + nit_avg * t_it,
        "unit": "s/iter",
        "cores": oracle.num_threads(),
        "kind": "port",
        "sample": ("oracle design iteration at %dx%dx%d from x0 with CG capped at 1 and 3 iterations "
                   "(%.1fs + %.1fs); per-CG-iteration %.3fs and fixed %.3fs extrapolated to the GPU's "
                   "mean %.1f CG iterations" % (args.nelx, args.nely, args.nelz, times[1], times[3], t_it,
                                                 t_fixed, nit_avg)),
    }


def probe_operator(args):
    """Short program for rocprofv3 --pmc passes: the level-0 operator at the
    bench size, then a plain device-to-device copy of known size (HIP's copy
    kernel) that calibrates FETCH_SIZE / WRITE_SIZE."""
    import gputo
    Ke, F, fixed = gputo.cantilever(args.nelx, args.nely, args.nelz)
    x0 = gputo.problems.perturbed_density(args.nelx, args.nely, args.nelz, 0.12, seed=args.seed)
    CX = 1e-9 + x0 ** 3 * (1 - 1e-9)
    s = gputo.Solver(args.nelx, args.nely, args.nelz, args.nl, precision=1)
    s.setup(Ke, CX, fixed)
    avg = ctypes.c_double()
    for k in [int(v) for v in args.probe_op.split(",")]:
        rc = s.lib.gputo_bench_operator(s.h, k, 3, ctypes.byref(avg))
        log("level-0 kernel %d: %.1f us (rc %d)" % (k, avg.value * 1e3, rc))
    hip = ctypes.CDLL("libamdhip64.so")
    n = 256 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    for _ in range(2):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    log("copy calibration: 2 copies of %d bytes" % n)
    hip.hipFree(a)
    hip.hipFree(b)
    s.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nelx", type=int, default=192)
    ap.add_argument("--nely", type=int, default=96)
    ap.add_argument("--nelz", type=int, default=96)
    ap.add_argument("--nl", type=int, default=5)
    ap.add_argument("--cg-tol", type=float, default=1e-8)
    ap.add_argument("--cg-maxit", type=int, default=300)
    ap.add_argument("--variants", default="f64,mixed_mfma3s1116,mixed_mfma3s1116+learned")
    ap.add_argument("--op-reps", type=int, default=50)
    ap.add_argument("--learn-steps", type=int, default=300)
    ap.add_argument("--learn-lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--probe-op", default="",
                    help="counter probe: a few launches of these level-0 kernels (ids as gputo_bench_operator) "
                         "and two 256 MiB device copies, then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.probe_op:
        probe_operator(args)
        return

    import gputo
    lib = gputo.load()
    Ke, F, fixed = gputo.cantilever(args.nelx, args.nely, args.nelz)
    x0 = gputo.problems.perturbed_density(args.nelx, args.nely, args.nelz, 0.12, seed=args.seed)
    nel = args.nelx * args.nely * args.nelz
    ndof = F.size

    variants = args.variants.split(",")
    res = {}
    for v in variants:
        res[v] = run_variant(gputo, v, args, Ke, F, fixed, x0)
    head = variants[-1]
    r = res[head]
    base = res.get("f64")

    # roofline of the dominant kernel (the level-0 operator of the head
    # variant).  The node-gather kernels (0: FP64, 1: FP32) run their 1152
    # flop per element on the vector ALUs, an arithmetic intensity above the
    # ridge of both precisions: priced against the compute peak of their
    # type.  The MFMA kernels (2-5) have the element products on the matrix
    # cores at a small fraction of their peak: priced against HBM.
    kernel = VARIANTS[head.replace("+learned", "")][1]
    s_bytes = 8 if kernel == 0 else 4
    op_bytes = (2 * ndof + nel) * s_bytes
    op_flops = 1152.0 * nel
    op_s = r["op_ms"] * 1e-3 if r["op_ms"] else None
    default_size = (args.nelx, args.nely, args.nelz, args.nl) == (192, 96, 96, 5)
    traffic = pmc_traffic(kernel) if default_size else None
    if kernel <= 1:
        roof = {"bound": "mfma", "achieved": op_flops / op_s / 1e12 if op_s else None,
                "peak": FP64_PEAK_TFLOPS if kernel == 0 else FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "traffic": traffic}
    else:
        roof = {"bound": "hbm", "achieved": op_bytes / op_s / 1e9 if op_s else None, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "traffic": traffic}
    roof["frac"] = roof["achieved"] / roof["peak"] if roof["achieved"] else None

    # the CPU baseline runs the FP64 reference algorithm: price it at the FP64 CG count
    nit_avg = float(np.mean((base or r)["cg_iters"]))
    out = {
        "metric": METRIC,
        "value": r["s_per_iter"],
        "unit": "s/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["s_per_iter"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "f64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": "cantilever %dx%dx%d hex (%d elements, %d DOFs), nl=%d, cg_tol %g, x0=0.12+U(-0.01,0.01) "
                        "seed %d, density filter rmin=sqrt(3), OC move 0.2, preset ref_c"
                        % (args.nelx, args.nely, args.nelz, nel, ndof, args.nl, args.cg_tol, args.seed),
            "variant": head,
            "fp32_level01_sweeps": level0_sweeps(head),
        },
        "roofline": roof,
        "matvec": {
            "kernel": "level-0 operator y = K(E) u (%s)" % head,
            "ms": r["op_ms"],
            "gelem_per_s": nel / op_s / 1e9 if op_s else None,
            "hbm_gbps_algorithmic": op_bytes / op_s / 1e9 if op_s else None,
            "hbm_frac": op_bytes / op_s / 1e9 / HBM_PEAK_GBS if op_s else None,
            "algorithmic_bytes": op_bytes,
            "algorithmic_flops": op_flops,
        },
        "cg_iters": r["cg_iters"],
        "solve_s_per_iter": r["solve_s_per_iter"],
        "variants": {v: {k: res[v][k] for k in ("s_per_iter", "solve_s_per_iter", "cg_iters", "op_ms",
                                               "learned_accepted", "learned_rejected", "train_s")}
                     for v in variants},
    }
    if base is not None and head != "f64":
        out["speedup_vs_f64"] = base["s_per_iter"] / r["s_per_iter"]
        Ub, Uh = base["U"], r["U"]
        out["u_rel_err_vs_f64"] = float(np.linalg.norm(Uh - Ub) / np.linalg.norm(Ub))
        out["compliance_rel_err_vs_f64"] = float(abs(r["compliance"][-1] - base["compliance"][-1])
                                                 / abs(base["compliance"][-1]))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args, Ke, F, fixed, x0, nit_avg)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
