"""The C ABI library loads and exports every entry point its headers declare (CPU only; no
compute call needs a GPU here)."""
import ctypes
import glob
import os
import re
import subprocess

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
