"""The C-ABI library loads and exports every entry point its header
declares (no compute call: this runs without a GPU)."""

import ctypes
import os
import re

from pfvs_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "pfvs_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b((?:pfg|pfa)_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_reference_interface():
    names = declared_functions()
    for ref in ["pfg_coni_batch", "pfg_output_free", "pfg_max_h11", "pfg_backend",
                "pfg_device_count", "pfg_release"]:
        assert ref in names


def test_library_exports_every_declared_symbol():
    path = gpu.lib_path()
    assert path and os.path.exists(path), "libpfvs_amd.so is not built"
    lib = ctypes.CDLL(path)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_static_facts_without_device():
    lib = gpu.require()
    assert lib.pfg_max_h11() == 41
    assert lib.pfg_backend() == b"hip"
    assert lib.pfa_arch() == b"gfx950"
