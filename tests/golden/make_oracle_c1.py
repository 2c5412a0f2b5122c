"""Write tests/golden/oracle_c1.npz: CG glsc3 histories of the oracle on
the full 32x32x32-element N=9 mesh (32,768,000 points), 100 iterations,
RHS c * dssum(U[-1,1) seed 0), brick geometry:

  jac_seq    Jacobi, Dirichlet mask, sequential dot products (glsc3 as
             src/math.f:603-617 sums them)
  jac_exact  the same solve with correctly rounded dot products
             (orc_glsc3_exact: compensated sums of the same rounded terms)
  ref_exact  identity preconditioner, the reference's mask list, correctly
             rounded dot products: the configuration of ref_c1.npz, whose
             history is the flang-compiled reference's own

The script also checks that the oracle's sequential identity/reference-
mask history equals ref_c1.npz bit for bit (the oracle pinned against the
reference at full size), so the exact histories differ from the
reference's only by the summation order of the dot products.

    OMP_NUM_THREADS=8 python tests/golden/make_oracle_c1.py   (~5 min)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from conftest import splitmix_uniform  # noqa: E402
from oracle import oracle as O  # noqa: E402


def history(mask, precond, exact, f=None):
    P = O.Problem(32, 32, 32, 10)
    P.set_mask_mode(mask)
    if precond == O.PRECOND_JACOBI:
        P.setup_jacobi()
    P.set_exact_dots(exact)
    if f is None:
        f = P.rhs_from_raw(splitmix_uniform(P.n, 0))
    _, _, h = P.cg(f, 100, precond)
    print("done", mask, precond, exact, flush=True)
    return h, f


def main():
    out = {}
    out["jac_seq"], f = history(O.MASK_DIRICHLET, O.PRECOND_JACOBI, 0)
    out["jac_exact"], _ = history(O.MASK_DIRICHLET, O.PRECOND_JACOBI, 1, f)
    out["ref_exact"], _ = history(O.MASK_REF, O.PRECOND_IDENTITY, 1, f)
    ref_seq, _ = history(O.MASK_REF, O.PRECOND_IDENTITY, 0, f)
    with np.load(os.path.join(HERE, "ref_c1.npz"), allow_pickle=False) as z:
        ref = z["cg_b323232_n10_hist"][0]
    if not np.array_equal(ref_seq, ref):
        raise SystemExit("oracle's sequential history differs from the reference's: %.3e" %
                         np.max(np.abs(ref_seq / ref - 1)))
    for k in ("ref_exact",):
        print(k, "vs reference: max rel %.3e" % np.max(np.abs(out[k] / ref - 1)))
    print("jac_seq vs jac_exact: max rel %.3e" %
          np.max(np.abs(out["jac_seq"] / out["jac_exact"] - 1)))
    np.savez_compressed(os.path.join(HERE, "oracle_c1.npz"), **out)


if __name__ == "__main__":
    main()
