"""The C-ABI library loads and exports every entry point declared in include/yasps_amd.h
(no compute calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

from yasps_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "yasps_amd.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(ysp_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "ysp_pcg_solve" in names and "ysp_get_indices" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built (python -m yasps_amd.build)")
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_load_attaches_signatures():
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = _lib.load()
    assert lib.ysp_version() >= 1


def test_one_hip_runtime_in_process():
    """The kernels and PyTorch must share one HIP runtime (one libamdhip64 mapped)."""
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    _lib.load()
    maps = Path("/proc/self/maps").read_text()
    runtimes = sorted(set(re.findall(r"/\S*libamdhip64\S*", maps)))
    assert len(runtimes) == 1, runtimes
