"""The C ABI: libhipvr.so loads without a GPU and exports every function that
include/hipvr.h declares; the ctypes mirror of hvr_params has the C layout."""

import ctypes
import os
import re
import subprocess

import pytest

from pyvr_amd.hip_renderer import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hipvr.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(hvr_[a-z0-9_]+)\s*\(", text))


def test_header_and_binding_agree():
    assert declared_functions() == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libhipvr.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(L, n)]
    assert not missing, missing
    L.hvr_abi_version.restype = ctypes.c_int
    assert L.hvr_abi_version() == 1


def test_loaded_binding_has_signatures():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libhipvr.so not built")
    L = _lib.lib()
    for name in _lib.EXPORTS:
        assert getattr(L, name).restype is not None


def test_params_struct_layout_matches_c(tmp_path):
    src = tmp_path / "layout.c"
    fields = [f for f, _ in _lib.HvrParams._fields_]
    body = "\n".join(f'  printf("{f} %zu\\n", offsetof(hvr_params, {f}));' for f in fields)
    src.write_text(f'#include <stdio.h>\n#include <stddef.h>\n#include "{HEADER}"\n'
                   f'int main(void) {{\n  printf("sizeof %zu\\n", sizeof(hvr_params));\n{body}\n  return 0;\n}}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), str(src)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    assert int(out["sizeof"]) == ctypes.sizeof(_lib.HvrParams)
    for f in fields:
        assert int(out[f]) == getattr(_lib.HvrParams, f).offset, f


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(RuntimeError, match="not found"):
        _lib.lib()
