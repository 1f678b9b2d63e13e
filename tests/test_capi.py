"""The C ABI library: it loads without a GPU and exports every entry point
include/gpmood.h declares (no compute calls here)."""
import ctypes
import os
import re

import numpy as np

from conftest import ROOT, load_package


def header_functions():
    text = open(os.path.join(ROOT, "include", "gpmood.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(gpm_[a-z_0-9]+)\s*\(", text, re.M)))


def test_header_declares_the_binding():
    gpm = load_package()
    assert header_functions() == sorted(gpm.capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    gpm = load_package()
    lib = ctypes.CDLL(gpm.capi.LIB_PATH)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of gpm_config / gpm_stats have the C compiler's layout."""
    import shutil
    import subprocess
    gpm = load_package()
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        import pytest
        pytest.skip("no C compiler")
    fields = {"gpm_config": gpm.capi.GpmConfig, "gpm_stats": gpm.capi.GpmStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{os.path.join(ROOT, "include", "gpmood.h")}"',
             "int main(void) {"]
    for cname, cls in fields.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        cname, what, val = line.split()
        cls = fields[cname]
        got = ctypes.sizeof(cls) if what == "size" else getattr(cls, what).offset
        assert got == int(val), line


def test_product_path_has_no_oracle_dependency():
    """The product package never imports or links the oracle."""
    pkg = os.path.join(ROOT, "gp-mood_amd")
    for dirpath, _, files in os.walk(pkg):
        if "_build" in dirpath:
            continue
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h", ".f90")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower() or f == "build.py", f


def test_run_noalloc_rejects_grids_beyond_its_scratch(gpm):
    """gpm_run_noalloc refuses a grid larger than its static scratch memory
    (2^20 cells) with GPM_EINVAL before any device call (runs without a GPU)."""
    import ctypes
    L = gpm.capi.load()
    c = gpm.Config(lf=2048, Mord=5, ngp=3).to_c()
    t = {k: np.zeros(4000, dtype=np.int32 if k == "ixiy" else np.float64) for k in ("zT", "d2x", "d2y", "gw", "ixiy")}
    u = np.zeros(16)
    rc = L.gpm_run_noalloc(ctypes.byref(c), *(gpm.capi.ptr(t[k]) for k in ("zT", "d2x", "d2y", "gw", "ixiy")),
                           gpm.capi.ptr(u), 1, 0.0, 1.0, None, None)
    assert rc == -1
