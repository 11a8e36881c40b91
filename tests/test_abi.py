"""C ABI checks that need no GPU: the library loads and exports every entry point that
include/fwi2d_hip.h declares, and the product never touches the test oracle."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "fwi2d_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fwi2d_[a-zA-Z0-9_]+)\s*\(", src)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for required in ("fwi2d_op_apply_forward", "fwi2d_op_apply_jacobianT", "fwi2d_engine_forward",
                     "fwi2d_engine_adjoint", "fwi2d_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    path = os.path.join(ROOT, "fwi2d_amd", "libfwi2d_hip.so")
    if not os.path.exists(path):
        pytest.skip("library not built")
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    from fwi2d_amd import lib as binding
    assert set(binding.EXPORTS) <= set(declared_functions())


def test_product_does_not_use_the_oracle():
    pkg = os.path.join(ROOT, "fwi2d_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower(), os.path.join(dirpath, f)
