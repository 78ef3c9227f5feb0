"""The C ABI library loads and exports every entry point its headers declare (CPU only; no
compute call needs a GPU here)."""
import ctypes
import glob
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(gr_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_functions():
    names = declared_functions()
    assert len(names) > 50, names


def test_library_exports_every_declared_symbol(gr):
    lib = ctypes.CDLL(os.path.join(ROOT, "graphite_amd", "libgraphite_hip.so"))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "graphite_amd", "libgraphite_hip.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\b(gr_\w+)\b", out))
    assert set(declared_functions()) <= exported


def test_python_binding_covers_the_abi(gr):
    from graphite_amd._capi import SIGNATURES
    assert set(declared_functions()) == set(SIGNATURES)


def test_abi_version_and_errors_without_device(gr):
    L = gr.load()
    assert L.gr_abi_version() == 1
    # host-only entry points work anywhere
    m = gr.Mesh.build(gr.Elements.box(1, 1, 1))
    assert m.sizes().n_cells == 1
    if gr.device_count() == 0:
        # the product path has no CPU fallback: device work fails loudly
        with pytest.raises(gr.GraphiteError) as ei:
            m.compute_metrics()
        assert ei.value.code == 5
        assert "HIP" in str(ei.value) or "device" in str(ei.value)


def test_shim_header_compiles():
    """The C++ header with the reference's names (cuMesh, FvmField, FvmEuler*) is valid C++."""
    shim = os.path.join(ROOT, "include", "graphite", "graphite.hpp")
    src = f'#include "{shim}"\nint main() {{ return 0; }}\n'
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-x", "c++", "-"], input=src, text=True,
                       capture_output=True)
    assert r.returncode == 0, r.stderr


def test_struct_layouts_match_the_header(gr, tmp_path):
    """The ctypes mirrors of the header's structs have the C layout (size and field offsets)."""
    from graphite_amd import _capi
    header = os.path.join(ROOT, "include", "graphite_hip.h")
    src = f"""#include <stddef.h>
#include <stdio.h>
#include "{header}"
int main(void) {{
    printf("%zu %zu %zu\\n", sizeof(gr_warm_norm), offsetof(gr_warm_norm, out_std), offsetof(gr_warm_norm, trust));
    printf("%zu\\n", sizeof(gr_wing_params));
    return 0;
}}
"""
    c = tmp_path / "layout.c"
    c.write_text(src)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", str(c), "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    W = _capi.WarmNorm
    assert [int(v) for v in out[:3]] == [ctypes.sizeof(W), W.out_std.offset, W.trust.offset]
    assert int(out[3]) == ctypes.sizeof(_capi.WingParams)
    n = gr.warm_norm(np.zeros((4, 6)), np.zeros((4, 5)))
    assert list(n.trust) == [2.0, 4.0]
