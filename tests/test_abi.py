"""The C ABI: liblfr.so loads on the CPU-only host and exports every entry
point declared in include/lfr.h; the ctypes table covers them all.  No
compute call is made here (no GPU)."""
import ctypes as C
import re

from conftest import ROOT

from lfric_core_amd import _ffi

HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names.update(re.findall(r"^\s*(?:int|int64_t|const char\*|void\*)\s+(lfr_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared()
    assert "lfr_matrix_vector" in names and "lfr_init" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(_ffi.LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_header():
    missing = [n for n in declared() if n not in _ffi.SIGNATURES]
    assert not missing, missing


def test_init_without_gpu_reports_error():
    import torch
    if torch.cuda.is_available():
        import pytest
        pytest.skip("GPU present")
    h = C.c_void_p()
    rc = _ffi.lib().lfr_init(0, C.byref(h))
    assert rc != 0 and not h.value
