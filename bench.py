#!/usr/bin/env python3
"""Benchmark of the CG Poisson hot path on one MI355X.

Workload (one "step"): one plain-FP64 Jacobi-preconditioned CG solve of
100 iterations on a 32x32x32-element box at N=9 (nx1=10, 32,768,000 local
points), brick geometry, homogeneous Dirichlet faces, RHS U[-1,1) seed 0
-> dssum -> *c.  Inputs are generated on the device before the timed
region.  value = local DOF x iterations / second (DOF = E*(N+1)^3, the
reference's definition, src/prox_dssum.f:32).

Also reported: the dominant kernel's roofline (element operator, HIP
events on the library stream during the timed steps), time to a 1e-10
relative residual on the deformed-geometry variant, and a CPU baseline
(the compiled reference when oracle/_ref was built, else the oracle) on a
bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import atexit
import json
import os
import re
import resource
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def cpu_baseline(nel, nx1, niter_sample):
    """Time the CPU path on a bounded sample: the same mesh, fewer
    iterations.  Returns the cpu_baseline record."""
    n = nel[0] * nel[1] * nel[2] * nx1 ** 3
    ref = os.path.join(ROOT, "oracle", "_ref", "c1", "refharness")
    if os.path.exists(ref) and nx1 <= 10 and nel[0] * nel[1] * nel[2] <= 32768:
        import numpy as np
        # the RHS file and the harness outputs are ~260 MB each at this size:
        # keep them out of the results folder and remove them afterwards
        tmp = tempfile.mkdtemp(prefix="nb_cpu_")
        atexit.register(shutil.rmtree, tmp, True)
        from oracle import oracle as O
        O.rand_uniform(n, 0).tofile(os.path.join(tmp, "f.bin"))

        def pre():
            resource.setrlimit(resource.RLIMIT_STACK,
                               (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

        inp = "bench\n%d %d %d %d %d 1\n%s\n%s\n" % (nel[0], nel[1], nel[2], nx1, niter_sample,
                                                    os.path.join(tmp, "f.bin"),
                                                    os.path.join(tmp, "o"))
        # the harness prints a line per finished setup phase; pass those on
        proc = subprocess.Popen([ref], stdin=subprocess.PIPE, stdout=subprocess.PIPE,
                                preexec_fn=pre, text=True)
        proc.stdin.write(inp)
        proc.stdin.close()
        lines = []
        for line in proc.stdout:
            lines.append(line)
            if "HK" in line:
                print("cpu_baseline:" + line.rstrip(), file=sys.stderr, flush=True)
        rc = proc.wait(timeout=900)
        m = re.search(r"HK cg call\s+1 niter\s+(\d+) n\s+(\d+) seconds\s+(\S+)", "".join(lines))
        if rc == 0 and m:
            secs = float(m.group(3))
            del np
            return {"value": n * niter_sample / secs, "unit": "DOF*iter/s", "cores": 1,
                    "kind": "reference",
                    "sample": "flang-compiled reference cg (src/cg.f, identity M^-1, its own "
                              "mask) on the full %dx%dx%d N=%d mesh, %d iterations, %.1f s" %
                              (nel[0], nel[1], nel[2], nx1 - 1, niter_sample, secs)}
    return cpu_all_cores(nel, nx1, niter_sample)


def cpu_reference_ranks(nel, nx1, niter_sample, ranks):
    """The flang-compiled reference on `ranks` host cores at once: one
    process per core, each on its slab of the mesh (nelx x nely x nelz/ranks
    elements, as an MPI decomposition along z gives it), without the halo
    exchange between slabs (the build has no MPI), all started together.
    Throughput = total points x iterations / the slowest process's time of
    the timed cg call, an upper bound for the reference on these cores."""
    ref = os.path.join(ROOT, "oracle", "_ref", "c1", "refharness")
    if not os.path.exists(ref) or nel[2] % ranks:
        return None
    sub = (nel[0], nel[1], nel[2] // ranks)
    n = sub[0] * sub[1] * sub[2] * nx1 ** 3
    tmp = tempfile.mkdtemp(prefix="nb_cpur_")
    atexit.register(shutil.rmtree, tmp, True)
    from oracle import oracle as O
    O.rand_uniform(n, 0).tofile(os.path.join(tmp, "f.bin"))

    def pre():
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    env = dict(os.environ, OMP_NUM_THREADS="1")
    procs = []
    for k in range(ranks):
        inp = "bench\n%d %d %d %d %d 1\n%s\n%s\n" % (sub[0], sub[1], sub[2], nx1, niter_sample,
                                                    os.path.join(tmp, "f.bin"),
                                                    os.path.join(tmp, "o%d" % k))
        p = subprocess.Popen([ref], stdin=subprocess.PIPE, stdout=subprocess.PIPE,
                             preexec_fn=pre, text=True, env=env)
        p.stdin.write(inp)
        p.stdin.close()
        procs.append(p)
    secs = []
    for p in procs:
        out = p.stdout.read()
        rc = p.wait(timeout=900)
        m = re.search(r"HK cg call\s+1 niter\s+(\d+) n\s+(\d+) seconds\s+(\S+)", out)
        if rc != 0 or not m:
            return None
        secs.append(float(m.group(3)))
    print("cpu_baseline: %d reference processes, cg seconds %.2f..%.2f" %
          (ranks, min(secs), max(secs)), file=sys.stderr, flush=True)
    return {"value": ranks * n * niter_sample / max(secs), "unit": "DOF*iter/s", "cores": ranks,
            "kind": "reference",
            "sample": "%d flang-compiled reference processes at once, one per core, each cg "
                      "(identity M^-1, its own mask) on a %dx%dx%d-element slab of the "
                      "%dx%dx%d N=%d mesh (no halo exchange), %d iterations, slowest %.1f s" %
                      (ranks, sub[0], sub[1], sub[2], nel[0], nel[1], nel[2], nx1 - 1,
                       niter_sample, max(secs))}


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_all_cores(nel, nx1, niter_sample):
    """The GPU line's own workload on every host core: the oracle's FP64
    Jacobi CG (Dirichlet faces, the same RHS) with OpenMP over elements in
    ax and over points in the vector updates, and its dot products summed
    per thread (compensated) and combined, on a bounded sample of
    iterations on the full mesh."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    Q = O.Problem(*nel, nx1)
    Q.set_mask_mode(O.MASK_DIRICHLET)
    Q.setup_jacobi()
    Q.set_exact_dots(1)
    f = Q.rhs_from_raw(O.rand_uniform(Q.n, 0))
    Q.cg(f, 1, O.PRECOND_JACOBI)  # warm-up: pages touched once
    t0 = time.perf_counter()
    Q.cg(f, niter_sample, O.PRECOND_JACOBI)
    secs = time.perf_counter() - t0
    return {"value": Q.n * niter_sample / secs, "unit": "DOF*iter/s", "cores": threads,
            "kind": "port",
            "sample": "oracle FP64 Jacobi CG (the GPU line's workload) on the full %dx%dx%d N=%d "
                      "mesh, OpenMP on %d threads, %d iterations, %.1f s" %
                      (nel[0], nel[1], nel[2], nx1 - 1, threads, niter_sample, secs)}


def Q_rtr0(Q, fq):
    """glsc3(r0, c, r0) of r0 = mask(f): the denominator of the relative
    residual (a zero-iteration tolerance solve reports it)."""
    import nekbone_amd as nb
    x, r = Q.vec(), Q.vec()
    _, rtr0 = Q.cg_tol_d(x, fq, r, 1e-300, 0, nb.PRECOND_JACOBI, maxiter_ok=True)
    x.free()
    r.free()
    return rtr0


def splitmix_u01(n, seed):
    """n values U[0,1) of the SplitMix64 stream with the given seed
    (value k = mix(seed + (k+1)*0x9E3779B97F4A7C15), top 53 bits)."""
    import numpy as np
    k = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def batched_workload(nb, count, seed=2024, graph=True):
    """count problems: E_k = round(exp(U[ln 64, ln 8192])) elements in the
    box cubic(E_k), N_k uniform in {5, 7, 9, 11}, brick geometry,
    Dirichlet faces, RHS U[-1,1) seed k -> dssum -> *c, Jacobi to 1e-8.
    Returns the batched and the one-at-a-time solve records."""
    import math
    u = splitmix_u01(2 * count, seed)
    probs, fs, xs = [], [], []
    for k in range(count):
        E = int(round(math.exp(math.log(64) + u[2 * k] * (math.log(8192) - math.log(64)))))
        nx1 = (6, 8, 10, 12)[min(3, int(4 * u[2 * k + 1]))]
        P = nb.Problem(*nb.cubic(E), nx1)
        P.set_mask_mode(nb.MASK_DIRICHLET)
        P.setup_jacobi()
        f, x = P.vec(), P.vec()
        P.rand_uniform_d(f, k)
        P.rhs_from_raw_d(f)
        probs.append(P), fs.append(f), xs.append(x)
    nb.sync()
    B = nb.Batch(probs)
    B.set_graph(graph)
    for P in probs:
        P.set_graph(graph)
    B.autotune(nb.PRECOND_JACOBI)
    path = B.path()
    tol, maxit = 1e-8, 5000
    B.cg_tol_d(xs, fs, [None] * count, tol, maxit)
    nb.sync()
    t0 = time.perf_counter()
    its, _ = B.cg_tol_d(xs, fs, [None] * count, tol, maxit)
    nb.sync()
    tb = time.perf_counter() - t0
    rs = [P.vec() for P in probs]
    nb.sync()
    t0 = time.perf_counter()
    its1 = []
    for P, f, x, r in zip(probs, fs, xs, rs):
        it, _ = P.cg_tol_d(x, f, r, tol, maxit, nb.PRECOND_JACOBI)
        its1.append(it)
    nb.sync()
    ts = time.perf_counter() - t0
    npts = sum(P.n for P in probs)
    work = sum(P.n * it for P, it in zip(probs, its))
    B.close()
    return {"problems": count, "local_points": npts, "tol": tol,
            "orders": sorted(set(P.nx1 - 1 for P in probs)),
            "iters_min": min(its), "iters_max": max(its), "same_iters_as_single": its == its1,
            "element_kernel_path": ("work queue, tile set %d" % path[0]) if path[0] >= 0
            else "one launch per order group", "x_update_in_element_kernel": bool(path[1]),
            "batched_ms": tb * 1e3, "one_at_a_time_ms": ts * 1e3, "speedup": ts / tb,
            "DOF_iter_per_s": work / tb}


def measured_traffic(kname, mesh):
    """HBM bytes per launch of kernel `kname` on `mesh`, as recorded in the
    counter summary under profiles/ (a line
    "traffic <kernel> <mesh> <bytes>"), or None if not recorded."""
    path = os.path.join(ROOT, "profiles", "ax_cg_hbm_traffic_32x32x32_n9.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.strip().strip("`").split()
        if len(f) == 4 and f[0] == "traffic" and f[1] == kname and f[2] == mesh:
            return float(f[3])
    return None


def tol_workload(nb, nel, nx1, delta):
    """Time to a 1e-10 relative residual on the deformed mesh (vertices
    moved by U[-0.1,0.1) of the element width, seed 1): plain FP64 Jacobi
    CG against the mixed-precision CG (FP32 inner iterations, FP64
    reliable updates, the same FP64 stopping rule)."""
    import numpy as np
    Q = nb.Problem(*nel, nx1)
    Q.set_mask_mode(nb.MASK_DIRICHLET)
    Q.set_geometry_deformed(0.1, 1)
    Q.setup_jacobi()
    fq, xq, rq = Q.vec(), Q.vec(), Q.vec()
    Q.rand_uniform_d(fq, 0)
    Q.rhs_from_raw_d(fq)
    def best_of(solve, reps=3):
        # the fastest of `reps` identical solves (each starts from x = 0)
        best, res = None, None
        for _ in range(reps):
            nb.sync()
            t0 = time.perf_counter()
            res = solve()
            nb.sync()
            ms = (time.perf_counter() - t0) * 1e3
            best = ms if best is None else min(best, ms)
        return best, res

    Q.cg_tol_d(xq, fq, rq, 1e-10, 5000, nb.PRECOND_JACOBI)
    fp64_ms, (iters, _) = best_of(lambda: Q.cg_tol_d(xq, fq, rq, 1e-10, 5000,
                                                     nb.PRECOND_JACOBI))
    x64 = xq.download()
    # the first call builds the FP32 copies; the autotuner picks the FP32
    # element-kernel shape (as nekbone_autotune does for the FP64 one)
    xm = Q.vec()
    Q.cg_mixed_d(xm, fq, rq, 1e-10, 5000, nb.PRECOND_JACOBI, delta)
    Q.autotune_mixed(nb.PRECOND_JACOBI)
    mixed_ms, (m_it, m_upd, m_rtr) = best_of(lambda: Q.cg_mixed_d(xm, fq, rq, 1e-10, 5000,
                                                                  nb.PRECOND_JACOBI, delta))
    xmh = xm.download()
    d = xmh - x64
    rtr0 = max(Q_rtr0(Q, fq), 1e-300)
    out = {
        "fp64_ms": fp64_ms, "fp64_iters": iters, "mixed_ms": mixed_ms,
        "mixed_iters": m_it, "mixed_fp64_updates": m_upd, "speedup": fp64_ms / mixed_ms,
        "mixed_final_rel_residual": (m_rtr / rtr0) ** 0.5,
        "mixed_vs_fp64_solution": {
            "l2_rel": float(np.linalg.norm(d) / np.linalg.norm(x64)),
            "max_rel": float(np.abs(d).max() / np.abs(x64).max())},
        "geometry": "deformed +-10% vertices, seed 1", "delta": delta or 1e-3,
        "timing": "fastest of 3 identical solves each, after a warm-up solve"}
    # the same solve with the FP16 factors, and with the FP16 matrix-core
    # inner operator
    out["mixed_variant_f32"] = list(Q.variant_f32())
    for name, mode in (("mixed_g16", nb.MIXED_G16), ("mixed_f16_mfma", nb.MIXED_F16_MFMA)):
        Q.set_mixed_inner(mode)
        Q.cg_mixed_d(xm, fq, rq, 1e-10, 5000, nb.PRECOND_JACOBI, delta)
        Q.autotune_mixed(nb.PRECOND_JACOBI)
        ms2, (it2, upd2, rtr2) = best_of(lambda: Q.cg_mixed_d(xm, fq, rq, 1e-10, 5000,
                                                              nb.PRECOND_JACOBI, delta))
        d2 = xm.download() - x64
        out[name] = {"ms": ms2, "iters": it2, "fp64_updates": upd2,
                     "variant_f32": list(Q.variant_f32()),
                     "speedup_vs_fp64": fp64_ms / ms2,
                     "final_rel_residual": (rtr2 / rtr0) ** 0.5,
                     "vs_fp64_solution": {
                         "l2_rel": float(np.linalg.norm(d2) / np.linalg.norm(x64)),
                         "max_rel": float(np.abs(d2).max() / np.abs(x64).max())}}
    Q.set_mixed_inner(nb.MIXED_F32)
    best = min((("mixed_f32", out["mixed_ms"]),) +
               tuple((k, out[k]["ms"]) for k in ("mixed_g16", "mixed_f16_mfma")),
               key=lambda kv: kv[1])
    out["best_mixed"] = {"inner": best[0], "ms": best[1], "speedup": fp64_ms / best[1]}
    return out


def hsmg_workload(nb, tol=1e-10, graph=True):
    """HSMG-preconditioned FP64 CG on a 16x16x16-element box at N=7,
    deformed +-10% (seed 1), Dirichlet faces, RHS U[-1,1) seed 0, to a
    1e-10 relative residual: with the FP64 smoother (the reference's
    arithmetic), with the FP32 local solves, each with the reference's
    coarse step and with the FP64 direct coarse solve, and Jacobi CG
    beside them."""
    import numpy as np
    P = nb.Problem(16, 16, 16, 8)
    P.set_mask_mode(nb.MASK_DIRICHLET)
    P.set_geometry_deformed(0.1, 1)
    P.set_graph(graph)
    t0 = time.perf_counter()
    P.setup_hsmg()
    setup_ms = (time.perf_counter() - t0) * 1e3
    P.setup_jacobi()
    f, x, r = P.vec(), P.vec(), P.vec()
    P.rand_uniform_d(f, 0)
    P.rhs_from_raw_d(f)
    out = {"elements": P.nelt, "nx1": 8, "local_points": P.n, "tol": tol,
           "geometry": "deformed +-10% vertices, seed 1", "setup_ms": setup_ms}
    xs = {}
    for name, pre, fp32, coarse in (("fp64_smoother", nb.PRECOND_HSMG, 0, 0),
                                    ("fp32_smoother", nb.PRECOND_HSMG, 1, 0),
                                    ("fp64_smoother_direct_coarse", nb.PRECOND_HSMG, 0, 1),
                                    ("fp32_smoother_direct_coarse", nb.PRECOND_HSMG, 1, 1),
                                    ("jacobi", nb.PRECOND_JACOBI, 0, 0)):
        P.set_hsmg_fp32(fp32)
        if coarse and "coarse_setup_ms" not in out:
            t0 = time.perf_counter()
            P.set_hsmg_coarse(1)
            out["coarse_setup_ms"] = (time.perf_counter() - t0) * 1e3
        P.cg_tol_d(x, f, r, tol, 5000, pre)
        nb.sync()
        t0 = time.perf_counter()
        it, _ = P.cg_tol_d(x, f, r, tol, 5000, pre)
        nb.sync()
        ms = (time.perf_counter() - t0) * 1e3
        out[name] = {"iters": it, "ms": ms, "ms_per_iter": ms / max(it, 1)}
        xs[name] = x.download()
    P.set_hsmg_fp32(0)
    P.set_hsmg_coarse(0)
    d = xs["fp32_smoother"] - xs["fp64_smoother"]
    ref = xs["fp64_smoother"]
    out["fp32_vs_fp64_smoother_solution"] = {
        "l2_rel": float(np.linalg.norm(d) / np.linalg.norm(ref)),
        "max_rel": float(np.abs(d).max() / np.abs(ref).max())}
    out["fp32_smoother_speedup"] = out["fp64_smoother"]["ms"] / out["fp32_smoother"]["ms"]
    d = xs["fp64_smoother_direct_coarse"] - xs["fp64_smoother"]
    out["direct_coarse_vs_reference_coarse_solution"] = {
        "l2_rel": float(np.linalg.norm(d) / np.linalg.norm(ref)),
        "max_rel": float(np.abs(d).max() / np.abs(ref).max())}
    out["direct_coarse_speedup"] = (out["fp64_smoother"]["ms"] /
                                    out["fp64_smoother_direct_coarse"]["ms"])
    for v in (f, x, r):
        v.free()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--niter", type=int, default=100)
    ap.add_argument("--nel", type=int, default=32)
    ap.add_argument("--nx1", type=int, default=10)
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--traffic", type=float, default=None,
                    help="HBM bytes per launch of the dominant kernel (default: the "
                         "figure recorded under profiles/ for this kernel and mesh)")
    ap.add_argument("--copy-probe", type=float, default=None, metavar="BYTES",
                    help="only run the copy kernel over BYTES (read+write) and print GB/s")
    ap.add_argument("--delta", type=float, default=0.0,
                    help="mixed precision: FP32 residual reduction between FP64 "
                         "residual updates (0: library default)")
    ap.add_argument("--batch", type=int, default=64,
                    help="problems in the batched-solve measurement (0: skip)")
    ap.add_argument("--no-hsmg", dest="hsmg", action="store_false",
                    help="skip the HSMG-preconditioned solve on 16x16x16 elements, N=7")
    ap.add_argument("--only", choices=("tol", "batch", "hsmg"), default=None,
                    help="run only this part (no timed CG steps) and print its record")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every kernel directly instead of replaying HIP graphs "
                         "(for tracing tools)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tol", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (one rank's share of the mesh); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import nekbone_amd as nb
    nb.init(0)
    if args.copy_probe:
        ms = nb.bw_probe(int(args.copy_probe), 10)
        print(json.dumps({"copy_bytes": int(args.copy_probe), "ms": ms,
                          "GBs": args.copy_probe / ms / 1e6}))
        return
    nel = (args.nel, args.nel, args.nel)
    nx1, niter = args.nx1, args.niter
    if args.only:
        part = {"tol": lambda: tol_workload(nb, nel, nx1, args.delta),
                "batch": lambda: batched_workload(nb, args.batch, graph=not args.no_graph),
                "hsmg": lambda: hsmg_workload(nb, graph=not args.no_graph)}[args.only]()
        print(json.dumps({args.only: part}))
        return
    P = nb.Problem(*nel, nx1)
    n = P.n
    P.set_graph(not args.no_graph)
    P.set_mask_mode(nb.MASK_DIRICHLET)
    P.setup_jacobi()
    f, x, r = P.vec(), P.vec(), P.vec()
    P.rand_uniform_d(f, 0)
    P.rhs_from_raw_d(f)
    nb.sync()
    print("bench: setup done, n=%d" % n, file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        P.cg_d(x, f, r, niter, nb.PRECOND_JACOBI)
    nb.sync()

    P.timing_enable(1 << nb.KCLASS_AX)
    P.timing_reset()
    nb.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        P.cg_d(x, f, r, niter, nb.PRECOND_JACOBI)
    nb.sync()
    t1 = time.perf_counter()
    launches, ax_ms = P.timing_get(nb.KCLASS_AX)
    P.timing_enable(0)
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = n * niter / ((t1 - t0) / args.steps)
    print("bench: %.3f ms/solve" % ms_step, file=sys.stderr, flush=True)

    # Dominant kernel: the element operator as launched inside CG, with the
    # direction update fused in (Jacobi in z form: the CG carries
    # z = M^-1 r).  Compulsory streams per local point: reads z, p (8 B
    # each), g (48 B), the point flag byte (1 B); writes p, w (8 B each);
    # with the deferred x update also reads and writes x (16 B).  With the
    # x-ring (the gather form's default) x is updated from the ring in
    # the update kernel instead and the element kernel is the variant
    # without it.
    epb, pf, xu = P.variant()
    gather, xring = P.cg_form()
    if xring:
        xu = 0
    bpp = 81 + (16 if xu else 0)
    kname = "k_ax<%d,%d,false,1,%d,%d>" % (nx1, epb, pf, xu)
    ax_avg_s = ax_ms / max(launches, 1) / 1e3
    bytes_fused = float(bpp) * n
    flops_ax = (19.0 + 12.0 * nx1) * nx1 ** 3 * P.nelt
    roofline = {"bound": "hbm", "achieved": bytes_fused / ax_avg_s / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": bytes_fused / ax_avg_s / 1e9 / HBM_PEAK_GBS,
                "traffic": None, "kernel": kname,
                "bytes_per_point": bpp, "avg_us": ax_avg_s * 1e6, "launches": launches,
                "gflops": flops_ax / ax_avg_s / 1e9,
                "fused": "ax_e + Jacobi p update (z form)" + (" + deferred x update" if xu else "")
                + (" + pap of shared copies (gather form)" if gather else ""),
                "cg_form": {"gather": gather, "x_ring": xring}}
    roofline["traffic"] = args.traffic if args.traffic is not None else \
        measured_traffic(kname, "%dx%dx%d" % nel)

    # The plain operator w = A_L u alone (64 B/point: u, 6 factors, w).
    P.timing_enable(1 << nb.KCLASS_AX)
    P.timing_reset()
    for _ in range(20):
        P.ax_d(r, f)
    nb.sync()
    la, ms_a = P.timing_get(nb.KCLASS_AX)
    P.timing_enable(0)
    plain_s = ms_a / max(la, 1) / 1e3
    ax_plain = {"avg_us": plain_s * 1e6, "bytes_per_point": 64,
                "achieved_GBs": 64.0 * n / plain_s / 1e9,
                "frac": 64.0 * n / plain_s / 1e9 / HBM_PEAK_GBS}

    extra = {}
    if not args.no_tol:
        print("bench: time to 1e-10", file=sys.stderr, flush=True)
        extra["time_to_1e-10"] = tol_workload(nb, nel, nx1, args.delta)
    if args.hsmg:
        print("bench: hsmg", file=sys.stderr, flush=True)
        extra["hsmg"] = hsmg_workload(nb, graph=not args.no_graph)

    if args.batch > 0:
        print("bench: batched solves", file=sys.stderr, flush=True)
        extra["batched"] = batched_workload(nb, args.batch, graph=not args.no_graph)

    cpu = None
    if not args.no_cpu:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(nel, nx1, args.cpu_iters)
        if cpu.get("kind") == "reference":
            print("bench: cpu baseline, all cores", file=sys.stderr, flush=True)
            ranks = min(int(os.environ.get("OMP_NUM_THREADS", host_cores())), host_cores())
            while ranks > 1 and nel[2] % ranks:
                ranks -= 1
            allc = cpu_reference_ranks(nel, nx1, args.cpu_iters, ranks) if ranks > 1 else None
            extra["cpu_baseline_all_cores"] = allc or cpu_all_cores(nel, nx1, args.cpu_iters)

    rec = {
        "metric": "CG DOF*iter/s + time to 1e-10 rel. residual, mixed vs FP64, N=9, 32^3 elems",
        "value": value, "unit": "DOF*iter/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "plain FP64 Jacobi CG, %d iterations, %dx%dx%d elements, N=%d, "
                               "brick, Dirichlet, RHS U[-1,1) seed 0" %
                               (niter, nel[0], nel[1], nel[2], nx1 - 1),
                   "elements": P.nelt, "nx1": nx1, "local_points": n, "iterations": niter,
                   "parallelism": "single GPU"},
        "roofline": roofline, "ax_plain": ax_plain, "cpu_baseline": cpu,
    }
    rec.update(extra)
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
