"""The Fortran side of the boundary: psb_hip_cbind_mod (iso_c_binding
interfaces to include/psb_hip.h) compiles with flang, and a Fortran driver
issuing the spasb(HLG) / spmm / gedot / krylov(CG) sequence through it runs
on the GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "psblas3_amd", "fortran")
DRIVER = os.path.join(ROOT, "psblas3_amd", "lib", "fortran", "hip_pde3d_driver")


@pytest.mark.skipif(shutil.which("flang") is None, reason="flang not installed")
def test_binding_module_compiles(tmp_path):
    for src in ("psb_hip_cbind_mod.F90", "hip_pde3d_driver.f90"):
        subprocess.check_call(["flang", "-fsyntax-only", "-module-dir", str(tmp_path),
                               os.path.join(FDIR, src)], cwd=tmp_path)


@pytest.mark.gpu
def test_fortran_driver_runs():
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (flang missing at build time)")
    print("case fortran driver", flush=True)
    out = subprocess.run([DRIVER], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0 and "PASS" in out.stdout
