"""The C-ABI libraries load and export every function their headers declare
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from wukonggpu_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wk[sge]_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("wukong_gstore.h", "libwkg_gstore.so"), ("wukong_gpu.h", "libwkg_hip.so"),
                                        ("wukong_engine.h", "libwkg_engine.so")])
def test_exports(header, lib):
    names = _declared(header)
    assert names, header
    if lib == "libwkg_engine.so":
        N.hip_lib()
    L = ctypes.CDLL(os.path.join(N.LIB_DIR, lib), mode=ctypes.RTLD_GLOBAL)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_bindings_load():
    N.gstore_lib()
    N.hip_lib()
    N.engine_lib()
