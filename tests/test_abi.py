"""The C-ABI library loads and exports every function include/*.h declares
(no device work: these run without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pairs_hip_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared()
    assert "pairs_hip_create" in names and "pairs_hip_run" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    from pairs_amd import native
    lib = native.lib()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    import importlib
    from pairs_amd import native
    for mod in ("dem", "batch", "mlp", "io"):  # each registers its own entry points
        if os.path.exists(os.path.join(ROOT, "pairs_amd", mod + ".py")):
            importlib.import_module("pairs_amd." + mod)
    assert set(_declared()) <= set(native.exported_symbols())


def test_product_path_has_no_oracle_import():
    """The product package never imports, loads or calls the CPU oracle."""
    for p in glob.glob(os.path.join(ROOT, "pairs_amd", "**", "*.py"), recursive=True):
        text = open(p).read()
        assert "oracle" not in re.sub(r"#.*", "", text).replace("oracle/", ""), p
