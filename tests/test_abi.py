"""The C-ABI library loads without a GPU and exports what include/volrus.h declares."""
import os
import re

from conftest import ROOT


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(vr_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_symbols():
    assert len(declared_symbols()) >= 20


def test_library_exports_every_declared_symbol(vr):
    lib = vr.load_library()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_header(vr):
    assert set(declared_symbols()) <= set(vr._SIGS), sorted(set(declared_symbols()) - set(vr._SIGS))


def test_no_cpu_fallback_in_product():
    """The product package never reaches the oracle."""
    pkg = os.path.join(ROOT, "sharc.volrus_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp", ".hpp")):
                text = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "oracle" not in text.lower(), f"{f} mentions the oracle"
