"""The Fortran iso_c_binding interface (lfric_core_amd/fortran) compiles with
flang, links against liblfr.so and calls through the C ABI."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FDIR = ROOT / "lfric_core_amd" / "fortran"
LIBDIR = ROOT / "lfric_core_amd" / "_lib"


def _build(tmp_path):
    if not (os.path.exists(FLANG) or shutil.which(FLANG)):
        pytest.skip("flang not available")
    exe = tmp_path / "lfr_binding_check"
    subprocess.run([FLANG, "-c", str(FDIR / "lfr_c_binding_mod.f90"), "-o", str(tmp_path / "b.o"),
                    "-module-dir", str(tmp_path)], check=True)
    subprocess.run([FLANG, str(FDIR / "lfr_binding_check.f90"), str(tmp_path / "b.o"),
                    "-module-dir", str(tmp_path), "-L", str(LIBDIR), "-llfr",
                    f"-Wl,-rpath,{LIBDIR}", "-o", str(exe)], check=True)
    return exe


def test_binding_compiles_links_and_reports_no_device(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present: covered by the gpu test")
    out = subprocess.run([str(_build(tmp_path))], capture_output=True, text=True, timeout=60)
    assert "lfr_init returned 2" in out.stdout      # LFR_ERR_HIP: no device


@pytest.mark.gpu
def test_binding_on_gpu(tmp_path):
    print("case fortran binding", flush=True)
    out = subprocess.run([str(_build(tmp_path))], capture_output=True, text=True, timeout=120)
    assert "lfr_space_create returned 0" in out.stdout, out.stdout + out.stderr
    assert "lfr_finalise returned 0" in out.stdout
