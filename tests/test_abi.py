"""CPU checks of the C-ABI boundary: the built library loads and exports every function the
public header declares (no device calls are made)."""
import ctypes as C
import os
import re

import pytest

from horsepower_amd import runtime
from horsepower_amd.vnode import V0

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "horsepower_amd.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*\**\s*(pfn\w+|hpa\w+)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


def test_header_declares_pfn_and_runtime():
    names = declared()
    assert "pfnCompress" in names and "pfnJoinIndex" in names and "hpaInit" in names
    assert len(names) > 60


def test_library_exports_every_declared_symbol():
    if not os.path.exists(runtime.LIB_PATH):
        pytest.fail("libhorsepower_amd.so is not built")
    dll = C.CDLL(runtime.LIB_PATH)
    missing = [n for n in declared() if not hasattr(dll, n)]
    assert not missing, missing


def test_node_layout_matches_reference():
    # src/horseir/v3/backend/common.h:31-55
    assert C.sizeof(V0) == 32
    assert V0.len.offset == 0 and V0.g.offset == 8 and V0.typ.offset == 16
