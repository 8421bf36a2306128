"""CPU check of the relink recipe: the reference's backend object keeps its
host-only primitives, its copies of the primitives this library replaces are made local with
objcopy and include/horsepower_amd.syms, and every other object then binds those names to
libhorsepower_amd.so.  A stand-in object plays h_primitive.o (no reference source is built);
nothing is run on a device."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SYMS = os.path.join(ROOT, "include", "horsepower_amd.syms")
LIBDIR = os.path.join(ROOT, "horsepower_amd")

# h_primitive.o's shape: a replaced primitive (pfnSum) and a host-only one (pfnTable) that
# calls it internally
STAND_IN = r"""
typedef struct V0 { long len; char *g; int typ; int pad; long u; } V0, *V;
int pfnSum(V z, V x){ (void)z; (void)x; return 42; }
int pfnTable(V z, V x){ return pfnSum(z, x); }
"""
CALLER = r"""
typedef struct V0 { long len; char *g; int typ; int pad; long u; } V0, *V;
int pfnSum(V z, V x);
int pfnTable(V z, V x);
int (*table[])(V, V) = { pfnSum, pfnTable };   /* the interpreter's dyaFunc-style table */
int main(void){ return table[0] == 0; }
"""


def declared_pfn():
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "horsepower_amd.h")).read(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_]\w*\s*\**\s*\**\s*(pfn\w+)\s*\(", text, flags=re.M)))


def test_symbol_list_matches_header():
    assert open(SYMS).read().split() == declared_pfn()


def test_localized_object_does_not_shadow_library(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libhorsepower_amd.so")):
        pytest.fail("libhorsepower_amd.so is not built")
    (tmp_path / "h_primitive.c").write_text(STAND_IN)
    (tmp_path / "main.c").write_text(CALLER)
    run = lambda *c: subprocess.run(list(c), cwd=tmp_path, check=True, capture_output=True, text=True).stdout
    run("gcc", "-c", "-O1", "h_primitive.c", "-o", "h_primitive.o")
    run("objcopy", f"--localize-symbols={SYMS}", "h_primitive.o", "h_primitive.local.o")
    local = run("nm", "h_primitive.local.o")
    assert re.search(r"\bt pfnSum\b", local) and re.search(r"\bT pfnTable\b", local)
    run("gcc", "-c", "main.c", "-o", "main.o")
    run("gcc", "main.o", "h_primitive.local.o", "-L", LIBDIR, "-lhorsepower_amd", "-o", "q")
    # the table's pfnSum is the library's (undefined here, bound at load time); the local copy
    # stays for the host-only primitive's own call
    dyn = run("nm", "-D", "q")
    assert re.search(r"\bU pfnSum\b", dyn), dyn
    assert not re.search(r"\bU pfnTable\b", dyn)
