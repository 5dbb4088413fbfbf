"""The C-ABI library loads without a GPU and exports every symbol include/dfr.h declares."""
import os
import re

import pytest

from difffr_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "dfr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dfr_[a-z_0-9]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(native.EXPORTS)


def test_library_exports_all_symbols():
    if not os.path.exists(native.LIB_PATH):
        pytest.fail("libdfr_hip.so is not built (run __graft_entry__.build())")
    assert native.missing_exports() == []


def test_struct_layouts_match_header(tmp_path):
    """sizeof/offsetof of dfr_params and dfr_stats as a C compiler sees them == ctypes mirror."""
    import ctypes as C
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dfr.h"', "int main(void){"]
    for st, cls in (("dfr_params", native.Params), ("dfr_stats", native.Stats)):
        lines.append(f'printf("{st} %zu\\n", sizeof({st}));')
        for name, _ in cls._fields_:
            lines.append(f'printf("{st}.{name} %zu\\n", offsetof({st}, {name}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                  text=True).stdout.splitlines())
    for st, cls in (("dfr_params", native.Params), ("dfr_stats", native.Stats)):
        assert int(out[st]) == C.sizeof(cls)
        for name, _ in cls._fields_:
            assert int(out[f"{st}.{name}"]) == getattr(cls, name).offset, (st, name)
