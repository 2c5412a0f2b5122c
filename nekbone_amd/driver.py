"""Host driver with the reference's run configuration and log format.

    python -m nekbone_amd.driver [data.rea] [--size SIZE] [--mgrid]
                                 [--precond identity|jacobi|hsmg]
                                 [--mask ref|dirichlet] [--rhs setf|random]

Mirrors the main program of the reference (src/driver.f:1-79): read the
run parameters from data.rea (read_param, :402-484), then for every
polynomial order nx1 = nx0..nxN step nxD and element count nelt =
iel0..ielN step ielD build the mesh (init_mesh, :117-194, one rank),
the right-hand side (set_f, :88-101), run cg twice and time the second
call (:54-60), and print the reference's lines: the two `cg:` lines of
each call (src/cg.f:44,77), the flop and time report
(set_timer_flop_cnt, :293-322) and the average MFlops (:66-78).  The
defaults reproduce the reference's default build: identity
preconditioner, the reference's own mask and set_f.  --mgrid (or
--precond hsmg) is the MGRID build, which sets nx0 >= 4 as read_param
does.  Every solve runs on the GPU through the HIP library.
"""
import argparse
import math
import re
import sys
import time

import numpy as np

from . import core as nb

FLOP_PER_POINT_CG = 15  # src/cg.f:81


class ParamError(Exception):
    """A data.rea the reference would reject (it prints and exits)."""


def _logical(tok):
    """Fortran list-directed LOGICAL input: optional '.', then T or F."""
    t = tok.strip().lstrip(".").upper()
    if t.startswith("T"):
        return True
    if t.startswith("F"):
        return False
    raise ParamError("ERROR READING    data.rea.....ABORT")


def read_size(path):
    """lx1 and lelt from a SIZE file (its `parameter (lx1=..)` and
    `parameter (lelt=..)` lines); None for an absent value."""
    txt = open(path).read().lower()
    out = {}
    for name in ("lx1", "lelt"):
        m = re.search(r"\b%s\s*=\s*(\d+)" % name, txt)
        out[name] = int(m.group(1)) if m else None
    return out


def read_param(text, lx1=None, lelt=None, mgrid=False, log=print):
    """read_param (src/driver.f:402-484) on the text of a data.rea: five
    lines, the values before each `=`.  Returns the dict of ifbrick,
    iel0, ielN, ielD, nx0, nxN, nxD, npx, npy, npz, mx, my, mz.  Raises
    ParamError where the reference aborts; the stride warning falls
    through to the nx0 < 4 message and sets nx0 = 4, as the reference's
    labels 220 and 230 do."""
    lines = [l for l in text.splitlines() if l.strip()]
    if len(lines) < 5:
        raise ParamError("ERROR READING    data.rea.....ABORT")
    try:
        vals = [l.split("=")[0].split() for l in lines[:5]]
        p = {"ifbrick": _logical(vals[0][0])}
        for names, v in zip((("iel0", "ielN", "ielD"), ("nx0", "nxN", "nxD"),
                             ("npx", "npy", "npz"), ("mx", "my", "mz")), vals[1:]):
            for name, s in zip(names, v[:3]):
                p[name] = int(s)
            if len(v) < 3:
                raise ValueError
    except (ValueError, IndexError):
        raise ParamError("ERROR READING    data.rea.....ABORT")
    if p["iel0"] > p["ielN"] or p["nx0"] > p["nxN"]:
        raise ParamError("ERROR data.rea :: iel0 > ielN or nx0 > nxN :: ABORT")
    if (lelt is not None and p["ielN"] > lelt) or (lx1 is not None and p["nxN"] > lx1):
        raise ParamError("ERROR data.rea : ielN>lelt or nxN>lx1(SIZE) :: ABORT")
    if p["ielD"] > p["ielN"] or p["nxD"] > p["nxN"]:
        log("WARNING data.rea : STRIDE   nxD>nxN or ielD>ielN !!!")
        log("ERROR data.rea nx0 must be greater than or equal to 4 ...setting nx0=4")
        p["nx0"] = 4
    elif p["nx0"] < 4 and mgrid:
        log("ERROR data.rea nx0 must be greater than or equal to 4 ...setting nx0=4")
        p["nx0"] = 4
    else:
        log(" ifmgrid    : %s    ifbrick    : %s" % ("T" if mgrid else "F",
                                                     "T" if p["ifbrick"] else "F"))
    return p


def init_mesh(p, nelt, log=print):
    """init_mesh (src/driver.f:117-194) on one rank: a chain of nelt
    elements, or a box of mx*my*mz (cubic(nelt) when the given triple does
    not multiply to nelt).  Returns (nelx, nely, nelz)."""
    if not p["ifbrick"]:
        npx, npy, npz, mx, my, mz = 1, 1, 1, nelt, 1, 1
    else:
        npx, npy, npz = p["npx"], p["npy"], p["npz"]
        if npx * npy * npz != 1:
            npx, npy, npz = nb.cubic(1)
        mx, my, mz = p["mx"], p["my"], p["mz"]
        if mx * my * mz != nelt:
            mx, my, mz = nb.cubic(nelt)
    nel = (mx * npx, my * npy, mz * npz)
    log("")
    log(" Processor Distribution:  npx,npy,npz= %11d %11d %11d" % (npx, npy, npz))
    log(" Element Distribution: nelx,nely,nelz= %11d %11d %11d" % nel)
    log(" Local Element Distribution: mx,my,mz= %11d %11d %11d" % (mx, my, mz))
    return nel


def _e(x):
    """Fortran 1pe12.4."""
    return "%12.4E" % x


def cg_lines(hist, niter):
    """The reference's two `cg:` lines (src/cg.f:44 and :77) from the
    glsc3 history [rtr0, (rtz, pap, rtr) per iteration]: iteration 0 with
    its residual norm, then iter = niter+1 with the last norm, alpha,
    beta and pap."""
    first = "cg:%4d%s" % (0, _e(math.sqrt(hist[0])))
    if niter == 0:
        return first, "cg:%4d%s%s%s%s" % (1, _e(math.sqrt(hist[0])), _e(0.0), _e(0.0), _e(0.0))
    rtz1, pap, rtr = hist[3 * niter - 2], hist[3 * niter - 1], hist[3 * niter]
    alpha = rtz1 / pap
    beta = 0.0 if niter == 1 else rtz1 / hist[3 * niter - 5]
    last = "cg:%4d%s%s%s%s" % (niter + 1, _e(math.sqrt(rtr)), _e(alpha), _e(beta), _e(pap))
    return first, last


def run_case(p, nx1, nelt, args, log=print):
    """One (nx1, nelt) case of the driver loop; returns its MFlops."""
    nel = init_mesh(p, nelt, log)
    P = nb.Problem(*nel, nx1)
    P.set_mask_mode(nb.MASK_REF if args.mask == "ref" else nb.MASK_DIRICHLET)
    precond = {"identity": nb.PRECOND_IDENTITY, "jacobi": nb.PRECOND_JACOBI,
               "hsmg": nb.PRECOND_HSMG}[args.precond]
    if precond == nb.PRECOND_JACOBI:
        P.setup_jacobi()
    if precond == nb.PRECOND_HSMG:
        P.setup_hsmg()
    niter = args.niter
    n = P.n
    f, x, r = P.vec(), P.vec(), P.vec()
    if args.rhs == "setf":
        P.set_f_d(f)
    else:
        P.rand_uniform_d(f, 0)
        P.rhs_from_raw_d(f)
    log("")
    hist = P.cg_d(x, f, r, niter, precond, want_hist=True)
    for line in cg_lines(hist, niter):
        log(line)
    nb.sync()
    t0 = time.perf_counter()
    hist = P.cg_d(x, f, r, niter, precond, want_hist=True)
    nb.sync()
    t1 = time.perf_counter() - t0
    for line in cg_lines(hist, niter):
        log(line)
    nxyz = nx1 ** 3
    flop_a = niter * float(19 * nxyz + 12 * nx1 * nxyz) * nelt  # src/cg.f:122
    flop_cg = (niter + 1) * FLOP_PER_POINT_CG * float(n)        # src/cg.f:81
    mflops = (flop_a + flop_cg) / (1e6 * t1) if t1 > 0 else 0.0
    log("")
    log("nelt = %7d, np = %9d, nx1 = %7d, elements =%10d" % (nelt, 1, nx1, nelt))
    log("Tot MFlops = %s, MFlops      = %s" % (_e(mflops), _e(mflops)))
    log("Setup Flop = %s, Solver Flop = %s" % (_e(flop_a), _e(flop_cg)))
    log("Solve Time = %s" % _e(t1))
    for v in (f, x, r):
        v.free()
    P.close()
    return mflops


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("rea", nargs="?", default="data.rea")
    ap.add_argument("--size", default=None, help="SIZE file: lx1 and lelt limits")
    ap.add_argument("--mgrid", action="store_true", help="the MGRID build (HSMG)")
    ap.add_argument("--precond", choices=("identity", "jacobi", "hsmg"), default=None)
    ap.add_argument("--mask", choices=("ref", "dirichlet"), default="ref")
    ap.add_argument("--rhs", choices=("setf", "random"), default="setf")
    ap.add_argument("--niter", type=int, default=100)
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args(argv)
    if args.precond is None:
        args.precond = "hsmg" if args.mgrid else "identity"
    mgrid = args.precond == "hsmg"
    lim = read_size(args.size) if args.size else {"lx1": None, "lelt": None}
    try:
        p = read_param(open(args.rea).read(), lim["lx1"], lim["lelt"], mgrid)
    except (OSError, ParamError) as e:
        print(str(e) if isinstance(e, ParamError) else "ERROR READING    data.rea.....ABORT")
        return 1
    nb.init(args.device)
    mfl = []
    for nx1 in range(p["nx0"], p["nxN"] + 1, max(1, p["nxD"])):
        for nelt in range(p["iel0"], p["ielN"] + 1, max(1, p["ielD"])):
            mfl.append(run_case(p, nx1, nelt, args))
    print("Avg MFlops = %s" % _e(sum(mfl) / len(mfl) if mfl else 0.0))
    sys.stdout.flush()
    return 0


if __name__ == "__main__":
    sys.exit(main())
