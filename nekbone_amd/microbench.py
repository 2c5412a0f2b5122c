"""Platform micro-benchmarks on one GPU: the element-operator contraction
sweep and the bandwidth probe, the one-GPU counterparts of the reference's
platform_timer (src/comm_mpi.f:541-553), whose mxm_test_all
(src/mxm_std.f:1107-1141) times the contraction shapes of ax over orders and
mxm variants to pick the fastest.

    python -m nekbone_amd.microbench [--points 16e6] [--orders 2-16]
        [--reps 5] [--json OUT] [--tune TUNING.json]

For every order nx1 and every compiled shape (elements per workgroup,
factor prefetch) of each inner arithmetic of the element operator
    f64      FP64 factors and contractions (the reference's ax_e)
    f32      FP32 factors and contractions
    g16      FP16 factors (per element and factor scale), FP32 contractions
    f16mfma  FP16 factors and FP16 matrix-core contractions (nx1 >= 4)
it times the operator kernel alone (HIP events on the library stream, the
library's per-kernel timers) on a box of about --points local points and
reports time per point, algorithmic HBM bytes per second (64, 32, 20 and 20
bytes per point) as a fraction of the 8 TB/s peak, and GFLOP/s of the
reference's flop model (19 + 12 nx1 per point, src/cg.f:122).  --tune also
runs the CG-iteration autotuner per order (FP64, Jacobi) and writes its
choices with nekbone_tuning_save, for later runs to load.
"""
import argparse
import json
import sys

import numpy as np

from . import core as nb

PEAK_GBS = 8000.0
BYTES = {"f64": 64, "f32": 32, "g16": 20, "f16mfma": 20}
INNER = {"f32": nb.MIXED_F32, "g16": nb.MIXED_G16, "f16mfma": nb.MIXED_F16_MFMA}


def _epb_default(nx1):
    return 16 if nx1 <= 3 else 8 if nx1 <= 4 else 4 if nx1 <= 5 else 2 if nx1 <= 8 else 1


def _orders(spec):
    lo, _, hi = spec.partition("-")
    return list(range(int(lo), int(hi or lo) + 1))


def _time_kernel(P, fn, reps):
    """Mean kernel time (ms) of the element-operator kernel over reps calls
    of fn (after one untimed call)."""
    fn()
    P.timing_reset()
    P.timing_enable(1 << nb.KCLASS_AX)
    for _ in range(reps):
        fn()
    n, ms = P.timing_get(nb.KCLASS_AX)
    P.timing_enable(0)
    return ms / max(n, 1)


def sweep_order(nx1, points, reps):
    """Rows for one order: every compiled shape of every arithmetic."""
    E = max(1, int(round(points / nx1 ** 3)))
    P = nb.Problem(*nb.cubic(E), nx1)
    P.set_mask_mode(nb.MASK_DIRICHLET)
    P.set_geometry_deformed(0.1, 1)
    rows = []
    n = P.n
    u64 = np.random.default_rng(nx1).uniform(-1.0, 1.0, n)
    u32 = u64.astype(np.float32)
    e0 = _epb_default(nx1)
    shapes = [(e, pf) for e in (e0, 2 * e0) for pf in (0, 1, 2, 3)]

    def row(kind, epb, pf, ms):
        us = ms * 1e3
        gbs = BYTES[kind] * n / (ms * 1e-3) / 1e9
        rows.append({"nx1": nx1, "kind": kind, "epb": epb, "pf": pf, "elements": P.nelt,
                     "points": n, "us": us, "ns_per_point": us * 1e3 / n, "GBs": gbs,
                     "frac_of_8TBs": gbs / PEAK_GBS,
                     "GFLOPs": (19 + 12 * nx1) * n / (ms * 1e-3) / 1e9})

    for epb, pf in shapes:
        try:
            P.set_variant(epb, pf, 1)
        except nb.NekboneError:
            continue
        if P.variant()[:2] != (epb, pf):
            continue
        row("f64", epb, pf, _time_kernel(P, lambda: P.ax_local(u64), reps))
    for kind, mode in INNER.items():
        try:
            P.set_mixed_inner(mode)
        except nb.NekboneError:
            continue
        for epb, pf in shapes:
            try:
                P.set_variant_f32(epb, pf, 1)
            except nb.NekboneError:
                continue
            row(kind, epb, pf, _time_kernel(P, lambda: P.ax_local_f32(u32), reps))
    P.close()
    return rows


def summarize(rows):
    """Fastest shape of each arithmetic per order, and where the matrix
    cores overtake the VALU kernels."""
    best = {}
    for r in rows:
        k = (r["nx1"], r["kind"])
        if k not in best or r["us"] < best[k]["us"]:
            best[k] = r
    orders = sorted({r["nx1"] for r in rows})
    table = []
    for nx1 in orders:
        ent = {"nx1": nx1}
        for kind in BYTES:
            b = best.get((nx1, kind))
            if b:
                ent[kind] = {"ns_per_point": round(b["ns_per_point"], 4), "epb": b["epb"],
                             "pf": b["pf"], "frac_of_8TBs": round(b["frac_of_8TBs"], 3)}
        table.append(ent)
    def faster_at(fast, slow):
        """Orders at which `fast` beats `slow` (fastest shape of each)."""
        return [nx1 for nx1 in orders if (nx1, fast) in best and (nx1, slow) in best and
                best[(nx1, fast)]["us"] < best[(nx1, slow)]["us"]]

    return {"best": table,
            "mfma_faster_than_f32_valu_at_nx1": faster_at("f16mfma", "f32"),
            "mfma_faster_than_fp16_factor_valu_at_nx1": faster_at("f16mfma", "g16")}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--points", type=float, default=16e6)
    ap.add_argument("--orders", default="2-16")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--json", default=None)
    ap.add_argument("--tune", default=None, metavar="TUNING.json")
    args = ap.parse_args(argv)
    nb.init(0)
    probe = nb.bw_probe(int(2 ** 30), 10)
    copy_gbs = 2 ** 30 / (probe * 1e-3) / 1e9
    print("copy probe: %.0f GB/s (%.3f of 8 TB/s)" % (copy_gbs, copy_gbs / PEAK_GBS), flush=True)
    rows = []
    for nx1 in _orders(args.orders):
        r = sweep_order(nx1, args.points, args.reps)
        rows += r
        for x in sorted(r, key=lambda x: (x["kind"], x["us"])):
            print("nx1 %2d %-8s epb %2d pf %d  %9.1f us  %.4f ns/pt  %6.0f GB/s  %.3f  %6.0f GF/s"
                  % (x["nx1"], x["kind"], x["epb"], x["pf"], x["us"], x["ns_per_point"],
                     x["GBs"], x["frac_of_8TBs"], x["GFLOPs"]), flush=True)
    out = {"copy_probe_GBs": copy_gbs, "rows": rows}
    out.update(summarize(rows))
    print(json.dumps({k: out[k] for k in ("best", "mfma_faster_than_f32_valu_at_nx1",
                                          "mfma_faster_than_fp16_factor_valu_at_nx1")}),
          flush=True)
    if args.tune:
        for nx1 in _orders(args.orders):
            E = max(1, int(round(args.points / nx1 ** 3)))
            P = nb.Problem(*nb.cubic(E), nx1)
            P.set_mask_mode(nb.MASK_DIRICHLET)
            P.setup_jacobi()
            ms = P.autotune(nb.PRECOND_JACOBI)
            print("autotune nx1 %2d: %s, %.3f ms per CG iteration" % (nx1, P.variant(), ms),
                  flush=True)
            P.close()
        nb.tuning_save(args.tune)
    if args.json:
        with open(args.json, "w") as f:
            json.dump(out, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
