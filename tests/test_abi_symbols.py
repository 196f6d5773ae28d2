"""The C-ABI library loads and exports every function include/*.h declares
(no GPU needed: nothing is called)."""
import glob
import os
import re

from tfce_mediation_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, re.M):
            if m.group(1) not in ("if", "defined"):
                names.add(m.group(1))
    return names


def test_library_exports_declared_symbols():
    lib = _lib.load_library()
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares a signature for each of them
    assert names <= set(_lib.SIGNATURES), sorted(names - set(_lib.SIGNATURES))
    assert lib.tfce_version().decode().startswith("tfcehip")
