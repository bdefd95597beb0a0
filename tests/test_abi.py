"""The C-ABI libraries load and export every function their headers declare
(no device calls: these run without a GPU)."""
import ctypes
import os
import re

from cgbdsm_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header, prefix):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z0-9_]+)\s*\(", text)))


def test_engine_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "cgbdsm_amd", "libcgbdsm.so"))
    names = _declared("cgbdsm.h", "cgbdsm_")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(N.DEV_SYMBOLS) <= set(names)


def test_host_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "cgbdsm_amd", "libcgbdsm_host.so"))
    names = _declared("cgbdsm_host.h", "cgh_")
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_the_header(tmp_path):
    """ctypes mirrors agree with the C compiler on every field offset."""
    structs = [("cgbdsm_plan", N.Plan), ("cgbdsm_stats", N.Stats), ("cgh_synth_spec", N.SynthSpec),
               ("cgh_batch", N.Batch), ("cgbdsm_split_stats", N.SplitStats)]
    rename = {"dele": "del"}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cgbdsm_host.h"', "int main(void){"]
    for cname, py in structs:
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {rename.get(f, f)}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    assert os.system(f"gcc -I{ROOT}/include {src} -o {exe}") == 0
    got = [int(x) for x in os.popen(str(exe)).read().split()]
    want = []
    for cname, py in structs:
        want.append(ctypes.sizeof(py))
        want.extend(getattr(py, f).offset for f, _ in py._fields_)
    assert got == want


def test_version_string():
    lib = N.dev_lib()
    assert b"gfx950" in lib.cgbdsm_version()
