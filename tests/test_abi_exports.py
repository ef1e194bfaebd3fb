"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", text):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.append(name)
    return names


def test_header_declares_the_reference_interface():
    names = declared_functions()
    for n in ("rhs_", "runge_kutta_", "update_state_", "tag_cells_", "rhs_hybrid_high_llf_", "rhs_drp_",
              "enl_last_error", "enl_set_stream", "enl_advance_patch", "enl_fill_physical_bc"):
        assert n in names
    assert len(names) == len(set(names))


def test_library_exports_every_declared_symbol():
    from enlightning_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libenlightning_hip.so is not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function has a ctypes prototype in the Python binding
    assert set(declared_functions()) <= set(_lib.PROTOTYPES)
    # the drop-in names are plain C symbols (Fortran mangling name_)
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for n in ("rhs_", "runge_kutta_", "update_state_", "tag_cells_"):
        assert n in syms


def test_no_cpu_fallback_in_product_path():
    # the package must not import the oracle anywhere
    for path in glob.glob(os.path.join(ROOT, "enlightning_amd", "**", "*.py"), recursive=True):
        text = open(path).read()
        assert "import oracle" not in text and "from oracle" not in text, path
