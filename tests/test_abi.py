"""The C ABI library builds for gfx950 and exports every entry point the
public header declares (no device needed: nothing is called)."""
import ctypes
import pathlib
import re

import ls1amd

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "ls1amd.h").read_text()
    return sorted(set(re.findall(r"\b(mdx_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(ls1amd.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(ls1amd.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_library_is_built_for_gfx950():
    data = ls1amd.LIB_PATH.read_bytes()
    assert b"gfx950" in data
