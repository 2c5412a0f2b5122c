"""Generate the golden fixtures under tests/golden/ from the compiled
reference (oracle/_ref/<variant>/refharness, built by `make -C oracle ref`
from the reference sources).  Run in a container that has the reference
tree; the fixtures are committed and the GPU tests only read them.

    python tests/golden/make_golden.py [--big]

Inputs are synthetic: U[-1,1) values from the counter-based splitmix64
stream (value k = mix(seed + (k+1)*0x9E3779B97F4A7C15)), generated here
with numpy so that no library is involved in making them.
"""
import argparse
import json
import os
import resource
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")


def splitmix_uniform(n, seed):
    """U[-1,1) stream shared by every component (see module doc)."""
    k = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return 2.0 * u - 1.0


def run(variant, task, nel, nx1, niter=0, ncalls=1, fin="-", out=None, env=None):
    exe = os.path.join(REF, variant, "refharness")
    if not os.path.exists(exe):
        raise SystemExit("missing %s: run `make -C oracle ref` first" % exe)

    def pre():
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    inp = "%s\n%d %d %d %d %d %d\n%s\n%s\n" % (task, nel[0], nel[1], nel[2], nx1, niter, ncalls,
                                              fin, out)
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run([exe], input=inp.encode(), capture_output=True, preexec_fn=pre, env=e)
    if r.returncode != 0:
        raise SystemExit("harness failed:\n" + r.stdout.decode() + r.stderr.decode())
    return r.stdout.decode()


def rd(prefix, ext, dtype=np.float64):
    return np.fromfile(prefix + "." + ext, dtype=dtype)


def small_cases(tmp):
    fx = {}
    # (variant, nel, nx1): ragged boxes, one element, several orders
    cases = [((2, 2, 2), 8), ((3, 2, 2), 10), ((1, 1, 1), 5), ((2, 3, 1), 12), ((4, 1, 1), 2),
             ((2, 1, 3), 6)]
    for nel, nx1 in cases:
        tag = "b%d%d%d_n%d" % (nel[0], nel[1], nel[2], nx1)
        pre = os.path.join(tmp, tag)
        run("id", "setup", nel, nx1, out=pre)
        fx[tag + "_dxm1"] = rd(pre, "dxm1")
        fx[tag + "_zgm1"] = rd(pre, "zgm1")
        fx[tag + "_wxm1"] = rd(pre, "wxm1")
        fx[tag + "_c"] = rd(pre, "c")
        fx[tag + "_g"] = rd(pre, "g")
        fx[tag + "_glo"] = rd(pre, "glo", np.int64)
        n = fx[tag + "_c"].size
        u = splitmix_uniform(n, 7)
        u.tofile(pre + ".in")
        run("id", "ax", nel, nx1, fin=pre + ".in", out=pre)
        fx[tag + "_ax_w1"] = rd(pre, "w1")
        fx[tag + "_ax_w2"] = rd(pre, "w2")
        cm = rd(pre, "cmask")
        fx[tag + "_cmask"] = cm[2:2 + int(cm[1])].astype(np.int32) - 1
        run("id", "dssum", nel, nx1, fin=pre + ".in", out=pre)
        fx[tag + "_dssum"] = rd(pre, "u")
    return fx


def cg_case(tmp, variant, nel, nx1, niter, ncalls, seed, tag):
    pre = os.path.join(tmp, tag)
    n = nel[0] * nel[1] * nel[2] * nx1 ** 3
    splitmix_uniform(n, seed).tofile(pre + ".in")
    histf = pre + ".hist"
    out = run(variant, "cg", nel, nx1, niter=niter, ncalls=ncalls, fin=pre + ".in", out=pre,
              env={"HK_GLSC3": histf})
    h = np.fromfile(histf).reshape(ncalls, 1 + 3 * niter)
    fx = {tag + "_hist": h}
    secs = [float(l.split()[-1]) for l in out.splitlines() if l.startswith("HK cg call")]
    return fx, secs, pre


def setf_fixture(tmp):
    """The reference's own right-hand side, set_f (src/driver.f:88-101),
    on 8x8x8 elements at N=7 (262,144 points, so i*i overflows the
    32-bit INTEGER from i = 46,341 on): a sample of the final f (after
    dssum and col2) at fixed indices, with the indices."""
    pre = os.path.join(tmp, "setf")
    run("id", "setf", (8, 8, 8), 8, out=pre)
    f = rd(pre, "f")
    idx = np.unique(np.concatenate([np.arange(0, 2048), np.arange(46300, 46400),
                                    np.arange(2048, f.size, 61), [f.size - 1]])).astype(np.int64)
    return {"setf_b888_n8_idx": idx, "setf_b888_n8_f": f[idx],
            "setf_b888_n8_sumsq": np.array([np.dot(f, f)])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--big", action="store_true", help="also the 32^3, N=9 history (minutes)")
    ap.add_argument("--setf", action="store_true",
                    help="only write ref_setf.npz (the set_f right-hand side)")
    args = ap.parse_args()
    if args.setf:
        with tempfile.TemporaryDirectory() as tmp:
            np.savez_compressed(os.path.join(HERE, "ref_setf.npz"), **setf_fixture(tmp))
        return
    meta = {"generator": "tests/golden/make_golden.py", "source": "oracle/_ref refharness"}
    with tempfile.TemporaryDirectory() as tmp:
        fx = small_cases(tmp)
        # CG, identity preconditioner, the reference's mask: 8^3 elements,
        # N=7, RHS U[-1,1) seed 0 -> dssum -> *c, two calls as the driver.
        f2, secs, pre = cg_case(tmp, "id", (8, 8, 8), 8, 100, 2, 0, "cg_b888_n8")
        fx.update(f2)
        fx["cg_b888_n8_x2_head"] = rd(pre, "x2")[:4096]
        fx["cg_b888_n8_r2_head"] = rd(pre, "r2")[:4096]
        x2 = rd(pre, "x2")
        fx["cg_b888_n8_x2_sumsq"] = np.array([np.dot(x2, x2)])
        meta["cg_b888_n8_seconds"] = secs
        f3, secs, _ = cg_case(tmp, "id", (3, 4, 5), 6, 40, 1, 3, "cg_b345_n6")
        fx.update(f3)
        # HSMG application and HSMG-preconditioned CG (MGRID build).
        pre = os.path.join(tmp, "mg")
        for nel, nx1, tag in [((2, 2, 2), 8, "mg_b222_n8"), ((3, 2, 2), 10, "mg_b322_n10"),
                              ((4, 4, 4), 8, "mg_b444_n8")]:
            n = nel[0] * nel[1] * nel[2] * nx1 ** 3
            splitmix_uniform(n, 11).tofile(pre + ".in")
            run("mg", "hsmg", nel, nx1, fin=pre + ".in", out=pre)
            fx[tag + "_z"] = rd(pre, "z")
            fx[tag + "_r"] = rd(pre, "r")
        f4, secs, _ = cg_case(tmp, "mg", (4, 4, 4), 8, 30, 1, 5, "mgcg_b444_n8")
        fx.update(f4)
        if args.big:
            f5, secs, _ = cg_case(tmp, "c1", (32, 32, 32), 10, 100, 1, 0, "cg_b323232_n10")
            fx.update(f5)
            meta["cg_b323232_n10_seconds"] = secs
            np.savez_compressed(os.path.join(HERE, "ref_c1.npz"), **f5)
    np.savez_compressed(os.path.join(HERE, "ref_small.npz"), **fx)
    with open(os.path.join(HERE, "ref_meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print("wrote", len(fx), "arrays")


if __name__ == "__main__":
    sys.exit(main())
