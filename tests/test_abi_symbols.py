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


def test_knobs_validate_their_arguments():
    """The process-wide knobs refuse out-of-range values with TFCE_EINVAL and
    a message (host-only calls, no GPU needed), and take valid ones."""
    lib = _lib.load_library()
    EINVAL = -1000
    assert lib.tfce_set_single_image_mode(2, 0) == EINVAL
    assert b"mode" in lib.tfce_last_error()
    assert lib.tfce_set_single_image_mode(-1, -3) == EINVAL
    assert lib.tfce_set_single_image_mode(-1, 0) == 0
    assert lib.tfce_set_occupancy(0) == EINVAL
    assert lib.tfce_ols_set_tile(4) == EINVAL
    assert lib.tfce_ols_set_tile(0) == 0
    assert lib.tfce_set_pair_sweep(1) == 0 and lib.tfce_set_lds_sweep(1) == 0
