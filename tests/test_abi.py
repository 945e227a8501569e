"""The C ABI: libhv.so loads and exports every entry point include/hv/hv.h declares."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

from helio_amd import _lib

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "hv" / "hv.h"


def declared_in_header() -> set[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(hv_[a-z0-9_]+)\s*\(", text))


def test_header_functions_are_exported():
    lib = C.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in sorted(declared_in_header()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert declared_in_header() <= set(_lib.declared_symbols()) | {"hv_last_error"}


def test_binding_loads_without_gpu():
    lib = _lib.load()
    assert lib.hv_mip_levels(1920, 1080) == 11
    assert b"" != lib.hv_last_error()


def test_no_oracle_in_product():
    """The product never imports, links or loads the oracle."""
    for p in list((ROOT / "helio_amd").rglob("*.py")) + list((ROOT / "helio_amd").rglob("*.hip")) + \
            list((ROOT / "helio_amd").rglob("*.cpp")) + list((ROOT / "helio_amd").rglob("*.h")):
        t = p.read_text()
        assert "import oracle" not in t and "hv_oracle" not in t and "or_cull" not in t, p
