"""The float32 acosf, cosf and square root that the GPU's excitation errors use
(ctemsoft_amd/csrc/ctem_sg.h) against the host C library that the reference's compiled
CalcsgSingle calls (diffraction.f90:649-674): every float input of the domains used, compiled
for the host with contraction off (tests/native/sg_check.cpp)."""
import os
import subprocess

from conftest import REPO


def test_sg_math_exhaustive(tmp_path):
    exe = str(tmp_path / "sg_check")
    src = os.path.join(REPO, "tests", "native", "sg_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                    "-Wno-unknown-pragmas", src, "-o", exe, "-lm"], check=True)
    env = dict(os.environ, OMP_NUM_THREADS=str(min(8, os.cpu_count() or 1)))
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "acosf 0 of 2130706434 differ" in r.stdout
    assert "cosf 0 of 1061752794 differ" in r.stdout


def test_libm_restatement_exhaustive(tmp_path):
    """expf, logf, sincosf and cabsf restated for the GPU's LUT build (ctemsoft_amd/csrc/ctem_libm.h)
    against the host C library that the reference's CalcUcg calls (diffraction.f90:515-626,
    others.f90:68-778), on every float of their domains (tests/native/libm_check.cpp). The FMA
    forms are checked, so the host needs the instruction."""
    with open("/proc/cpuinfo") as f:
        if " fma" not in f.read():
            import pytest
            pytest.skip("host processor without fused multiply-add: its C library takes the non-FMA build")
    exe = str(tmp_path / "libm_check")
    src = os.path.join(REPO, "tests", "native", "libm_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-mfma", "-fopenmp",
                    "-Wno-unknown-pragmas", src, "-o", exe, "-lm"], check=True)
    env = dict(os.environ, OMP_NUM_THREADS=str(min(8, os.cpu_count() or 1)))
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=900)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    for name in ("expf", "logf", "sincosf", "cabsf", "cexpf"):
        assert f"{name} 0 of" in r.stdout
