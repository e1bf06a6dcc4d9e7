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
# of the levels from 2 on[, sweeps of the levels from 3 on]]])
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
    "mixed_mfma3s2214": (2, 3, 2, 2, 0, 1, 1, 4, 4),  # 2 + 2 on level 0, 1 + 1 on level 1, 4 + 4 below
    "mixed_mfma3s2216": (2, 3, 2, 2, 0, 1, 1, 6, 6),  # 2 + 2 on level 0, 1 + 1 on level 1, 6 + 6 below
    "mixed_mfma3s2224": (2, 3, 2, 2, 0, 2, 2, 4, 4),  # 2 + 2 on levels 0 and 1, 4 + 4 below
    "mixed_mfma3s1314": (2, 3, 1, 3, 0, 1, 1, 4, 4),  # 1 + 3 on level 0, 1 + 1 on level 1, 4 + 4 below
    "mixed_mfma3s1312": (2, 3, 1, 3, 0, 1, 1, 2, 2),  # 1 + 3 on level 0, 1 + 1 on level 1, 2 + 2 below
    "mixed_mfma3s1316": (2, 3, 1, 3, 0, 1, 1, 6, 6),  # 1 + 3 on level 0, 1 + 1 on level 1, 6 + 6 below
    "mixed_mfma3s1311": (2, 3, 1, 3, 0, 1, 1, 1, 1),  # 1 + 3 on level 0, 1 + 1 on level 1, 1 + 1 below
    "mixed_mfma3s1313": (2, 3, 1, 3, 0, 1, 1, 3, 3),  # 1 + 3 on level 0, 1 + 1 on level 1, 3 + 3 below
    # 1 + 3 on level 0, 1 + 1 on level 1, 2 + 2 (1 + 1) on level 2, 8 + 8 (10 + 10) from level 3
    "mixed_mfma3s1312c8": (2, 3, 1, 3, 0, 1, 1, 2, 2, 8, 8),
    "mixed_mfma3s1312c10": (2, 3, 1, 3, 0, 1, 1, 2, 2, 10, 10),
    "mixed_mfma3s1311c8": (2, 3, 1, 3, 0, 1, 1, 1, 1, 8, 8),
    # as mixed_mfma3s1311c8 with level 1 pre-smoothed by the prerelaxed iterate alone (-1: no sweep)
    "mixed_mfma3s1311c8p1": (2, 3, 1, 3, 0, -1, 1, 1, 1, 8, 8),
    "mixed_mfma3s1334": (2, 3, 1, 3, 0, 1, 3, 4, 4),  # 1 + 3 on levels 0 and 1, 4 + 4 below
    "mixed_mfma3s1324": (2, 3, 1, 3, 0, 2, 2, 4, 4),  # 1 + 3 on level 0, 2 + 2 on level 1, 4 + 4 below
    "mixed_mfma3s3314": (2, 3, 3, 3, 0, 1, 1, 4, 4),  # 3 + 3 on level 0, 1 + 1 on level 1, 4 + 4 below
    "mixed_mfma3s1514": (2, 3, 1, 5, 0, 1, 1, 4, 4),  # 1 + 5 on level 0, 1 + 1 on level 1, 4 + 4 below
    "mixed_mfma3s2414": (2, 3, 2, 4, 0, 1, 1, 4, 4),  # 2 + 4 on level 0, 1 + 1 on level 1, 4 + 4 below
    "mixed_mfma3s1314l2": (2, 3, 1, 3, 2, 1, 1, 4, 4),  # 1314 with the previous coarsest factor in 2 V-cycles
    "mixed_mfma3s1314l3": (2, 3, 1, 3, 3, 1, 1, 4, 4),  # ... in 3 V-cycles
}


def variant_of(name):
    """The sweep/kernel tuple of a variant name; its suffixes are "+learned"
    (the learned level-2 stage, learned = 2) and "+graph" (the V-cycle
    replayed from a recorded HIP graph, graph_vcycles = 1)."""
    return VARIANTS[name.replace("+learned", "").replace("+graph", "")]


def level2_sweeps(v):
    """(pre, post) sweeps of the levels from 2 on (0, 0: the preset's)."""
    return (v[7], v[8]) if len(v) > 7 else (0, 0)


def level3_sweeps(v):
    """(pre, post) sweeps of the levels from 3 on (0, 0: those from 2 on)."""
    return (v[9], v[10]) if len(v) > 9 else (0, 0)


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
    v = variant_of(name)
    if v[0] == 0:
        return None
    def cnt(x, y=0):  # a sweep count (0: the next setting, else the preset's 3; -1: none)
        return max(x, 0) if x else (max(y, 0) if y else 3)

    n1 = level1_sweeps(v)
    s0 = "%d+%d" % (cnt(v[2]), cnt(v[3]))
    s1 = "%d+%d" % (cnt(n1[0]), cnt(n1[1]))
    out = s0 if s0 == s1 else "%s on level 0, %s on level 1" % (s0, s1)
    n2 = level2_sweeps(v)
    n3 = level3_sweeps(v)
    if n3[0] or n3[1]:
        out += ", %d+%d on level 2, %d+%d from level 3" % (cnt(n2[0]), cnt(n2[1]), cnt(n3[0], n2[0]),
                                                         cnt(n3[1], n2[1]))
    elif n2[0] or n2[1]:
        out += ", %d+%d on the levels from 2" % (cnt(n2[0]), cnt(n2[1]))
    return out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pretrain(gputo, s, args, Ke, fixed, name):
    """Training of the learned level-2 stage before the design loop, after the
    config-3 protocol: the network learns from the level-2 defects that plain
    MGCG solves of other cantilevers log (volume fraction ~ U[0.08, 0.3] with
    a +-0.01 uniform perturbation, a unit downward line load along j at a
    random height k of the free end face x = nelx); after each solve,
    `pretrain_steps` Adam steps on its logged defects with its own level-2
    operator.  A one-off cost, reported beside the design iterations."""
    rng = np.random.default_rng(args.seed + 1000)
    nel = args.nelx * args.nely * args.nelz
    nn = (args.nely + 1) * (args.nelz + 1)
    t0 = time.perf_counter()
    samples = 0
    loss = []
    for p in range(args.pretrain_problems):
        vf = rng.uniform(0.08, 0.3)
        x = np.clip(vf + rng.uniform(-0.01, 0.01, nel), 1e-3, 1.0)
        CX = 1e-9 + x ** 3 * (1 - 1e-9)
        k = int(rng.integers(0, args.nelz + 1))
        F = np.zeros(3 * (args.nelx + 1) * nn)
        nodes = args.nelx * nn + (args.nely + 1) * k + np.arange(args.nely + 1)
        F[3 * nodes + 2] = -1.0 / (args.nely + 1)
        U, r, n, rc = s.solve(Ke, F, CX, fixed, args.cg_tol, args.cg_maxit)
        if rc != 0:
            raise RuntimeError("pretraining solve failed: %d" % rc)
        samples += s.learn_info()["nsamples"]
        l0, l1 = s.learn_train(args.pretrain_steps, args.learn_lr)
        loss.append((l0, l1))
    s.learn_logging(False)
    s.lib.gputo_device_synchronize()
    dt = time.perf_counter() - t0
    log("[%s] pretraining: %d problems, %d logged level-2 defects, %d Adam steps, %.2fs" % (
        name, args.pretrain_problems, samples, args.pretrain_problems * args.pretrain_steps, dt))
    return {"seconds": dt, "problems": args.pretrain_problems, "samples": samples,
            "adam_steps": args.pretrain_problems * args.pretrain_steps,
            "energy_first_last": [loss[0][0], loss[-1][1]] if loss else None}


def run_variant(gputo, name, args, Ke, F, fixed, x0):
    """One design run of variant `name`: `warmup` untimed design iterations,
    `steps` timed ones (bracketed by device synchronisation), then untimed
    ones up to the `horizon` (each still timed by the library's clock).
    The learned variant trains after each warm-up iteration on the level-2
    defects its solve logged; that training time is counted into the timed
    steps' cost (`s_per_iter` includes it, amortised over them).  The state
    (xPhys, U, compliance) is kept at design iterations 1, 50 and 100."""
    v = variant_of(name)
    prec, kernel, npre0, npost0, lag = v[:5]
    npre1, npost1 = level1_sweeps(v)
    npre2, npost2 = level2_sweeps(v)
    npre3, npost3 = level3_sweeps(v)
    learned = "+learned" in name
    s = gputo.Solver(args.nelx, args.nely, args.nelz, args.nl, precision=prec, learned=2 if learned else 0,
                     l0_kernel=kernel, mixed_npre0=npre0, mixed_npost0=npost0, mixed_npre1=npre1,
                     mixed_npost1=npost1, mixed_npre2=npre2, mixed_npost2=npost2, coarse_lag=lag,
                     graph_vcycles="+graph" in name, mixed_npre3=npre3, mixed_npost3=npost3)
    pre = pretrain(gputo, s, args, Ke, fixed, name) if learned else None
    run = gputo.TopOpt(s, Ke, x0, F, fixed, volfrac=0.12, penal=3.0, rmin=np.sqrt(3.0), ft=1, move=0.2,
                       cg_tol=args.cg_tol, cg_maxit=args.cg_maxit)
    keep = {}

    def snap(it, st):
        if it + 1 in REPORT_ITERS:
            x, xp, U = run.state()
            keep[it + 1] = {"xPhys": xp, "U": U, "C": st["compliance"]}

    train_s = 0.0  # training inside the design loop (none: the network is trained beforehand)
    hist = []
    for w in range(args.warmup):
        st = run.iterate()
        hist.append(st)
        snap(len(hist) - 1, st)
        log("[%s] warmup %d: C=%.6e nit=%d solve=%.3fs total=%.3fs" % (name, w, st["compliance"], st["cg_iters"],
                                                                       st["seconds_solve"], st["seconds_total"]))
    s.lib.gputo_device_synchronize()
    t0 = time.perf_counter()
    timed = []
    for _ in range(args.steps):
        st = run.iterate()
        timed.append(st)
    s.lib.gputo_device_synchronize()
    wall = time.perf_counter() - t0
    for st in timed:
        hist.append(st)
        snap(len(hist) - 1, st)
    for k, st in enumerate(timed):
        log("[%s] step %d: C=%.8e nit=%d relres=%.2e solve=%.3fs total=%.3fs" % (
            name, k, st["compliance"], st["cg_iters"], st["relres"], st["seconds_solve"], st["seconds_total"]))
    while len(hist) < args.horizon:
        st = run.iterate()
        hist.append(st)
        snap(len(hist) - 1, st)
        if len(hist) % 10 == 0:
            log("[%s] design iteration %d: C=%.8e nit=%d total=%.3fs" % (name, len(hist), st["compliance"],
                                                                         st["cg_iters"], st["seconds_total"]))
    x, xp, U = run.state()
    if learned:
        on = [h for h in hist if h["learned_accepted"] + h["learned_rejected"] > 0]
        log("[%s] the learned stage ran in %d of %d design iterations" % (name, len(on), len(hist)))
    out = {
        "pretraining": pre,
        "learned_on_iters": [k + 1 for k, h in enumerate(hist) if h["learned_accepted"] + h["learned_rejected"] > 0],
        "s_per_iter": (wall + train_s) / args.steps,
        "s_per_iter_excl_training": wall / args.steps,
        "train_s": train_s,
        "solve_s_per_iter": float(np.mean([h["seconds_solve"] for h in timed])),
        "cg_iters": [h["cg_iters"] for h in timed],
        "compliance": [h["compliance"] for h in timed],
        "relres": [h["relres"] for h in timed],
        "learned_accepted": [h["learned_accepted"] for h in timed],
        "learned_rejected": [h["learned_rejected"] for h in timed],
        "U": U,
        "keep": keep,
        "warmup_s": [h["seconds_total"] for h in hist[:args.warmup]],
        "warmup_cg_iters": [h["cg_iters"] for h in hist[:args.warmup]],
    }
    if len(hist) > args.warmup + args.steps:
        out["horizon"] = {
            "design_iters": len(hist),
            "s_per_iter": float(np.mean([h["seconds_total"] for h in hist])),
            "cg_iters_total": int(sum(h["cg_iters"] for h in hist)),
            "cg_iters_every_10th": [h["cg_iters"] for h in hist[9::10]],
        }
    # dominant kernel: the level-0 operator, timed with HIP events on the
    # solver's stream on the hierarchy of the last solve
    avg = ctypes.c_double()
    rc = s.lib.gputo_bench_operator(s.h, kernel, args.op_reps, ctypes.byref(avg))
    out["op_ms"] = avg.value if rc == 0 else None
    run.close()
    s.close()
    return out


REPORT_ITERS = (1, 50, 100)


def fp64_errors(gputo, args, Ke, F, fixed, res, base):
    """Error of u and compliance of every reported result of a mixed variant
    against the FP64 plain MGCG solve of the same design (per result), and
    the drift of its design trajectory from the FP64 run's (trajectory)."""
    out = {}
    s = gputo.Solver(args.nelx, args.nely, args.nelz, args.nl)
    for it, k in sorted(res["keep"].items()):
        CX = 1e-9 + k["xPhys"] ** 3 * (1 - 1e-9)
        U64, r64, n64, rc = s.solve(Ke, F, CX, fixed, args.cg_tol, args.cg_maxit)
        U64[1::3] *= -1  # the MATLAB frame of gputo_mgcg1 back to the C frame of the loop state
        C64 = float(F @ U64)
        e = {"u_rel_err": float(np.linalg.norm(k["U"] - U64) / np.linalg.norm(U64)),
             "compliance_rel_err": abs(k["C"] - C64) / abs(C64), "fp64_cg_iters": int(n64)}
        if base is not None and it in base["keep"]:
            b = base["keep"][it]
            e["trajectory_compliance_rel_diff"] = abs(k["C"] - b["C"]) / abs(b["C"])
            e["trajectory_max_abs_dxPhys"] = float(np.abs(k["xPhys"] - b["xPhys"]).max())
        out[str(it)] = e
    s.close()
    return out


def physical_cores():
    """Physical cores of the host (lscpu's core,socket pairs), or None."""
    try:
        import subprocess
        txt = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        return len({l for l in txt.splitlines() if l and not l.startswith("#")}) or None
    except Exception:
        return None


def cpu_baseline(gputo, args, Ke, F, fixed, x0, base):
    """The oracle (C/OpenMP restatement of Cantilever/omp/mgcg1.c and the
    design steps of Cantilever/omp/top3D125mod.m) on the host cores, measured:
    the first `cpu_iters` design iterations of the 192x96x96 run from the same
    x0 (whole iterations, CG to the same tolerance), and the 20 design
    iterations of the 48x24x24 cantilever (config 0: 4 levels, the MATLAB
    omega 0.6 1+1 smoother, CG tolerance 1e-10) beside the GPU's FP64 run of
    the same 20 iterations.  Threads: the reference's 16 (mgcg1.c:120-121),
    fewer if the process may use fewer CPUs."""
    import oracle
    threads = min(16, len(os.sched_getaffinity(0)))
    oracle.set_num_threads(threads)
    st = oracle.DesignState(args.nelx, args.nely, args.nelz, x0)
    times, nits = [], []
    for it in range(args.cpu_iters):
        t0 = time.perf_counter()
        o = oracle.design_iter(st, Ke, F, fixed, args.nl, args.cg_tol, args.cg_maxit, 0.12, 3.0, np.sqrt(3.0), 1,
                               2.0, 0.2)
        times.append(time.perf_counter() - t0)
        nits.append(int(o["cgiters"]) + 1)
        log("[cpu] design iteration %d at %dx%dx%d: %.2fs, CG %d (rc %d)" % (it + 1, args.nelx, args.nely, args.nelz,
                                                                             times[-1], nits[-1], o["rc"]))
    # config 0 on both sides
    n0 = (48, 24, 24, 4)
    K0, F0, fx0 = gputo.cantilever(*n0[:3])
    xc0 = np.full(n0[0] * n0[1] * n0[2], 0.12)
    st0 = oracle.DesignState(*n0[:3], xc0)
    t0 = time.perf_counter()
    for it in range(20):
        o = oracle.design_iter(st0, K0, F0, fx0, n0[3], 1e-10, 500, 0.12, 3.0, np.sqrt(3.0), 1, 2.0, 0.2,
                               preset=oracle.PRESET_MATLAB_CFG0)
    cpu0 = (time.perf_counter() - t0) / 20
    s0 = gputo.Solver(*n0, preset=gputo.PRESET_MATLAB_CFG0)
    r0 = gputo.TopOpt(s0, K0, xc0, F0, fx0, volfrac=0.12, penal=3.0, rmin=np.sqrt(3.0), ft=1, move=0.2, cg_tol=1e-10,
                      cg_maxit=500)
    s0.lib.gputo_device_synchronize()
    t0 = time.perf_counter()
    for it in range(20):
        g = r0.iterate()
    s0.lib.gputo_device_synchronize()
    gpu0 = (time.perf_counter() - t0) / 20
    r0.close()
    s0.close()
    log("[cpu] 48x24x24 config 0: oracle %.3fs, GPU FP64 %.4fs per design iteration" % (cpu0, gpu0))
    n_same = min(args.cpu_iters, len(base["warmup_s"])) if base else 0
    return {
        "value": float(np.mean(times)),
        "unit": "s/iter",
        "cores": threads,
        "kind": "port",
        "sample": ("oracle (C/OpenMP restatement of mgcg1.c + top3D125mod.m), design iterations 1-%d of the "
                   "%dx%dx%d run from the same x0, each whole (CG %s to tol %g), %d threads; the GPU window's "
                   "designs need more CG iterations, so this is a lower bound on the CPU's time there"
                   % (args.cpu_iters, args.nelx, args.nely, args.nelz, "/".join(map(str, nits)), args.cg_tol,
                      threads)),
        "physical_cores": physical_cores(),
        "threads": threads,
        "design_s": times,
        "cg_iters": nits,
        "gpu_f64_same_designs_s_per_iter": float(np.mean(base["warmup_s"][:n_same])) if n_same else None,
        "config0_48x24x24": {"iters": 20, "cpu_s_per_iter": cpu0, "gpu_f64_s_per_iter": gpu0},
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
    hip = ctypes.CDLL("libamdhip64.so.7")
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
    ap.add_argument("--variants", default="f64,mixed_mfma3s1311c8p1,mixed_mfma3s1311c8p1+learned")
    ap.add_argument("--op-reps", type=int, default=50)
    ap.add_argument("--pretrain-problems", type=int, default=16)
    ap.add_argument("--pretrain-steps", type=int, default=50)
    ap.add_argument("--learn-lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=3, help="whole oracle design iterations at the bench size")
    ap.add_argument("--horizon", type=int, default=100,
                    help="design iterations per variant run (the timed steps are inside it; 0: only those)")
    ap.add_argument("--probe-op", default="",
                    help="counter probe: a few launches of these level-0 kernels (ids as gputo_bench_operator) "
                         "and two 256 MiB device copies, then exit")
    args = ap.parse_args()
    args.horizon = max(args.horizon, args.warmup + args.steps)
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
    kernel = variant_of(head)[1]
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
        "variants": {v: {k: res[v][k] for k in ("s_per_iter", "s_per_iter_excl_training", "solve_s_per_iter",
                                               "cg_iters", "op_ms", "learned_accepted", "learned_rejected",
                                               "train_s", "pretraining", "learned_on_iters", "horizon")
                         if k in res[v] and res[v][k] is not None}
                     for v in variants},
    }
    if base is not None and head != "f64":
        out["speedup_vs_f64"] = base["s_per_iter"] / r["s_per_iter"]
        Ub, Uh = base["U"], r["U"]
        out["u_rel_err_vs_f64"] = float(np.linalg.norm(Uh - Ub) / np.linalg.norm(Ub))
        out["compliance_rel_err_vs_f64"] = float(abs(r["compliance"][-1] - base["compliance"][-1])
                                                 / abs(base["compliance"][-1]))
        out["speedup_vs_f64_mixed_without_learned"] = {
            v: base["s_per_iter"] / res[v]["s_per_iter"] for v in variants if v not in ("f64", head)}
    # error of u and compliance against the FP64 plain solve of the same
    # design, for every reported result (design iterations 1, 50, 100)
    out["fp64_error_per_result"] = {v: fp64_errors(gputo, args, Ke, F, fixed, res[v], base)
                                    for v in variants if v != "f64"}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(gputo, args, Ke, F, fixed, x0, base)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
