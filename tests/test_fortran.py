"""The Fortran iso_c_binding module (tbfsolver_amd/fortran/tbfgpu_mod.f90) builds
against libtbfgpu.so, and its driver runs solveFPS and solvePCG on the GPU the
way the reference's poissonEqn calls them (fields with one halo layer)."""
import os
import subprocess

import pytest

from tbfsolver_amd import build as B


def test_fortran_binding_builds():
    if not os.path.exists(B.FLANG):
        pytest.skip("flang not available")
    exe = B.build_fortran()
    assert exe and os.path.exists(exe)
    nm = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True).stdout
    for sym in ("tbfg_fps_create", "tbfg_fps_solve", "tbfg_pcg_create", "tbfg_pcg_solve", "tbfg_last_error"):
        assert sym in nm, sym


@pytest.mark.gpu
def test_fortran_driver_on_gpu(torch_gpu):
    if not os.path.exists(B.FORTRAN_TEST):
        pytest.skip("Fortran driver not built")
    print("case fortran driver", flush=True)
    r = subprocess.run([B.FORTRAN_TEST], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "tbfgpu_test OK" in r.stdout
