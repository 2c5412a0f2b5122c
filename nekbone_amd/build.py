"""Build libnekbone_amd.so (HIP, gfx950) in-tree with hipcc.

    python -m nekbone_amd.build [-v]

Objects go to nekbone_amd/_build/, the library to nekbone_amd/lib/.
Every HIP source is compiled with -ffp-contract=off (no implicit FMA; the
kernels request fused multiply-adds explicitly where wanted) and with
packed-FP32 instructions disabled.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIBNAME = "libnekbone_amd.so"
ARCH = os.environ.get("NEKBONE_ARCH", "gfx950")

SOURCES = ["nb_ax_i10.hip", "nb_ax_i12.hip", "nb_ax_i8.hip", "nb_ax_i6.hip", "nb_ax_ilow.hip",
           "nb_ax_ihigh.hip", "nb_setup.cpp", "nb_ax.hip", "nb_kernels.hip", "nb_cg.hip",
           "nb_api.hip",
           "nb_legacy.hip",
           "nb_ax32_low.hip", "nb_ax32_high.hip", "nb_ax32_mf.hip", "nb_ax_ib.hip", "nb_ax_iq.hip", "nb_mixed.hip", "nb_batch.hip", "nb_hsmg.hip", "nb_hsmg_setup.cpp", "nb_hsmg_coarse.hip",
           "nb_noalloc.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(ROOT, "include")]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def lib_path():
    return os.path.join(LIBDIR, LIBNAME)


def _newest_header():
    hdrs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    hdrs += [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))]
    return max(os.path.getmtime(h) for h in hdrs)


def _compile(src, verbose, incremental=False):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src.rsplit(".", 1)[0] + ".o")
    if incremental and os.path.exists(obj):
        if os.path.getmtime(obj) > max(os.path.getmtime(path), _newest_header()):
            return obj
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip"] + COMMON + DEVICE + ["-c", path, "-o", obj]
    else:
        cmd = [HIPCC, "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"] + COMMON + [
            "-c", path, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (src, r.stdout, r.stderr))
    if verbose and r.stderr.strip():
        print(r.stderr, flush=True)
    return obj


def build(verbose=False, jobs=None, incremental=False):
    """Compile every source (incremental=True skips objects newer than
    their source and every header) and link the library."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose, incremental), srcs))
    tmp = lib_path() + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + [
        "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, lib_path())
    build_fortran_host(verbose)
    return lib_path()


FORTRAN = os.path.join(HERE, "fortran")
HOST = os.path.join(LIBDIR, "nekbone_host")


def flang():
    """The image's flang, or None."""
    for cand in (os.environ.get("FLANG"), "/opt/rocm/llvm/bin/flang", shutil.which("flang")):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran_host(verbose=False):
    """Link the Fortran 2003 example host (nekbone_amd/fortran) against the
    library: lib/nekbone_host, found next to libnekbone_amd.so at run time.
    Returns its path, or None when no flang is installed."""
    fc = flang()
    if fc is None:
        return None
    mod = os.path.join(BUILD, "fortran")
    os.makedirs(mod, exist_ok=True)
    cmd = [fc, "-O2", "-J", mod, "-o", HOST + ".tmp",
           os.path.join(FORTRAN, "nekbone_amd_if.f90"), os.path.join(FORTRAN, "nekbone_host.f90"),
           "-L" + LIBDIR, "-lnekbone_amd", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("fortran host build failed\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(HOST + ".tmp", HOST)
    return HOST


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))
