"""The C-ABI library loads and exports every symbol the headers declare (CPU)."""
import ctypes
import glob
import os
import re

import pytest

from tincr_amd import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(tda_\w+)\s*\(", text, flags=re.M):
            syms.add(m.group(1))
    return syms


def test_headers_declare_the_api():
    syms = declared_symbols()
    assert set(nat.EXPORTS) <= syms
    assert {"tda_debug_gemm", "tda_debug_eigh"} <= syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(nat.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(nat.LIB_PATH)
    missing = [s for s in sorted(declared_symbols()) if not hasattr(lib, s)]
    assert not missing, missing


def test_orbe_differences_host_entry():
    """tda_orbe_differences is host code: safe to call without a GPU."""
    if not os.path.exists(nat.LIB_PATH):
        pytest.skip("library not built")
    from tincr_amd.tda import orbe_differences
    import numpy as np
    d = orbe_differences([-1.0, -0.5], [0.1, 0.3, 0.7])
    assert np.array_equal(d, [1.1, 1.3, 1.7, 0.6, 0.8, 1.2])
