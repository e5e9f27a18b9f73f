"""The Fortran host side of the boundary: ctemsoft_amd/fortran/ctem_ebsd_iface.f90 (ISO_C_BINDING
interfaces) driven by ctem_lgh_check.f90, the same call a CTEMsoft host makes in place of
CalcLgh3 (CTEMEBSDmaster.f90:453)."""
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest
import torch

from conftest import REPO, load_golden

EXE = os.path.join(REPO, "ctemsoft_amd", "ctem_lgh_check")


def _write_input(d, fx, j, k):
    dm = np.asfortranarray(fx[f"dyn{j}_dynmat"])
    lam = fx["lam"].astype(np.float32)
    with open(os.path.join(d, "lgh_in.bin"), "wb") as f:
        f.write(struct.pack("<ii", dm.shape[0], len(lam)))
        f.write(struct.pack("<ddd", float(fx["thick"]), float(fx["kn"][k]), float(fx["dz"])))
        f.write(lam.tobytes())
        f.write(dm.T.astype(np.complex128).tobytes())  # column-major
    return dm.shape[0]


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error path")
def test_fortran_host_reports_library_error_without_gpu():
    if not os.path.exists(EXE):
        pytest.skip("Fortran check program not built")
    fx = load_golden("ni_c0")
    with tempfile.TemporaryDirectory() as d:
        _write_input(d, fx, 0, int(fx["dyn_ik"][0]))
        r = subprocess.run([EXE], cwd=d, capture_output=True, text=True)
        assert r.returncode == 1
        assert "ctem_ctx_create failed" in r.stdout and "no HIP device" in r.stdout


@pytest.mark.gpu
def test_fortran_host_calclgh3_on_gpu():
    assert os.path.exists(EXE), "build the Fortran check program with python -m ctemsoft_amd.build"
    fx = load_golden("fe_c1w")
    for j, k in enumerate(fx["dyn_ik"][:2]):
        with tempfile.TemporaryDirectory() as d:
            n = _write_input(d, fx, j, int(k))
            r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=120)
            print(r.stdout, flush=True)
            assert r.returncode == 0, r.stdout + r.stderr
            L = np.fromfile(os.path.join(d, "lgh_out.bin"), np.complex128).reshape(n, n).T
        Lr = fx[f"dyn{j}_lgh"]
        err = np.linalg.norm(L - Lr) / np.linalg.norm(Lr)
        print(f"Fortran host, direction {int(k)}: Lgh rel err {err:.2e}", flush=True)
        assert err < 1e-10
