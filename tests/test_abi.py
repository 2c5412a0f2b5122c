"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

from conftest import ROOT


def _declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)
        src = re.sub(r"//[^\n]*", " ", src)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", src):
            name = m.group(1)
            start = src.rfind(";", 0, m.start())
            start = max(start, src.rfind("}", 0, m.start()), src.rfind("{", 0, m.start()))
            decl = src[start + 1:m.start()]
            if "typedef" in decl or "#" in decl or name in ("defined",):
                continue
            # a declaration: return type words precede the name
            if re.search(r"\b(void|int|double|int64_t|const|char)\b[\s\*]*$", decl):
                names.add(name)
    return sorted(names)


def _lib_path():
    from nekbone_amd import core
    if not os.path.exists(core.LIB_PATH):
        subprocess.run(["python3", "-m", "nekbone_amd.build"], cwd=ROOT, check=True)
    return core.LIB_PATH


def test_header_declares_reference_entry_points():
    names = set(_declared_functions())
    for ref in ("cg_", "ax_", "ax_e_", "local_grad3_", "local_grad3_t_", "dssum_", "solvem_",
                "glsc3_", "gs_setup_", "gs_op_", "gs_free_", "maskit_"):
        assert ref in names, ref


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib_path())
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_no_cpu_fallback_in_product():
    """The product package never imports the oracle."""
    for path in glob.glob(os.path.join(ROOT, "nekbone_amd", "**", "*"), recursive=True):
        if path.endswith((".py", ".hip", ".cpp", ".h", ".f90", ".F90")):
            text = open(path, errors="replace").read()
            assert "oracle" not in re.sub(r"#.*|//.*", "", text).lower() or path.endswith(".md"), path
