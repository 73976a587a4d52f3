"""The Fortran iso_c_binding interface (lfric_core_amd/fortran) compiles with
flang, links against liblfr.so and calls through the C ABI: the Fortran
snippets of the integration guide compile as written against the module, and a
Fortran program drives a W2 mass apply (reference layout and packed), a CG
solve and a generic BiCGstab solve on the GPU, checked against the oracle."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FDIR = ROOT / "lfric_core_amd" / "fortran"
LIBDIR = ROOT / "lfric_core_amd" / "_lib"

# LFRic's log_mod for compiling the snippets (the snippets report errors
# through log_event as a PSy layer does)
LOG_STUB = """module log_mod
  implicit none
  integer, parameter :: LOG_LEVEL_ERROR = 4
contains
  subroutine log_event(msg, level)
    character(len=*), intent(in) :: msg
    integer, intent(in) :: level
    write(*, '(A)') msg
    if (level >= LOG_LEVEL_ERROR) error stop 1
  end subroutine log_event
end module log_mod
"""


def _flang():
    if not (os.path.exists(FLANG) or shutil.which(FLANG)):
        pytest.skip("flang not available")


def _module(tmp_path):
    _flang()
    subprocess.run([FLANG, "-c", str(FDIR / "lfr_c_binding_mod.f90"), "-o", str(tmp_path / "b.o"),
                    "-module-dir", str(tmp_path)], check=True)
    return tmp_path / "b.o"


def _build(tmp_path):
    obj = _module(tmp_path)
    exe = tmp_path / "lfr_binding_check"
    subprocess.run([FLANG, str(FDIR / "lfr_binding_check.f90"), str(obj),
                    "-module-dir", str(tmp_path), "-L", str(LIBDIR), "-llfr",
                    f"-Wl,-rpath,{LIBDIR}", "-o", str(exe)], check=True)
    return exe


def test_integration_snippets_compile(tmp_path):
    """Every ```fortran block of the integration guide is a complete program unit that
    compiles against lfr_c_binding_mod (with a stub of LFRic's log_mod)."""
    _module(tmp_path)
    text = (ROOT / "INTEGRATION.md").read_text()
    blocks = re.findall(r"```fortran\n(.*?)```", text, flags=re.S)
    assert len(blocks) >= 3
    src = tmp_path / "snippets.f90"
    src.write_text(LOG_STUB + "\n".join(blocks))
    out = subprocess.run([FLANG, "-c", str(src), "-module-dir", str(tmp_path), "-I", str(tmp_path),
                          "-o", str(tmp_path / "snippets.o")], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr


def test_binding_compiles_links_and_reports_no_device(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present: covered by the gpu test")
    out = subprocess.run([str(_build(tmp_path))], capture_output=True, text=True, timeout=60)
    assert "lfr_init returned 2" in out.stdout      # LFR_ERR_HIP: no device


def _write_problem(path, cfg):
    ncell, ndf = cfg.w2.dofmap.shape
    col = cfg.colouring
    cic = np.asfortranarray(col.cells_in_colour.astype(np.int32))
    hdr = np.array([ndf, ncell, cfg.nlayers, cfg.w2.undf, col.ncolours, cic.shape[1]], np.int32)
    x = np.linspace(-1.0, 1.0, cfg.w2.undf)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.ascontiguousarray(cfg.w2.dofmap, dtype=np.int32).tobytes())   # map(ndf, ncell)
        f.write(cic.tobytes(order="F"))
        f.write(np.ascontiguousarray(col.ncells_per_colour, dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(cfg.mass).tobytes())    # local_stencil(ncell_3d, ndf, ndf)
        f.write(x.tobytes())
        f.write(cfg.b.tobytes())
    return x


@pytest.mark.gpu
def test_binding_drives_the_gpu(tmp_path):
    import oracle
    from oracle.solvers import bicgstab_solve, cg_solve
    from lfric_core_amd.configs import precision_config
    print("case fortran binding W2 mass C4L6", flush=True)
    cfg = precision_config(4, 6, with_div_grad=False)
    x = _write_problem(tmp_path / "problem.bin", cfg)
    out = subprocess.run([str(_build(tmp_path)), str(tmp_path / "problem.bin"), str(tmp_path)],
                         capture_output=True, text=True, timeout=120)
    assert "lfr_finalise returned 0" in out.stdout, out.stdout + out.stderr
    ap = lambda v: oracle.invoke_coloured(0, cfg.colouring, cfg.nlayers, np.zeros_like(v), v, cfg.mass,  # noqa: E731
                                          cfg.w2.dofmap, cfg.w2.dofmap)
    ref = ap(x)
    load = lambda n: np.fromfile(tmp_path / n, dtype=np.float64)   # noqa: E731
    np.testing.assert_array_equal(load("y_ref.bin"), ref)
    np.testing.assert_array_equal(load("y_packed.bin"), ref)
    xo = np.zeros(cfg.w2.undf)
    cg_solve(ap, xo, cfg.b, max_iter=25, monitor=False)
    assert np.linalg.norm(load("x_cg.bin") - xo) <= 1e-11 * np.linalg.norm(xo)
    diag = oracle.mm_diagonal(np.zeros(cfg.w2.undf), cfg.mass, cfg.w2.dofmap, cfg.nlayers, cfg.colouring)
    xb = np.zeros(cfg.w2.undf)
    it, conv, _, _ = bicgstab_solve(ap, xb, cfg.b, prec=lambda v: v / diag, r_tol=1e-10, max_iter=200)
    assert f"bicgstab iterations {it} converged 1" in out.stdout, out.stdout
    assert np.linalg.norm(load("x_bicg.bin") - xb) <= 1e-10 * np.linalg.norm(xb)
