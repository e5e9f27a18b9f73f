"""EBSD detector patterns (include/ctem_detector.h), CPU side.

The fixture tests/golden/ebsd_det.npz holds the reference's detector-pattern program
(src_programs/CTEMEBSD.f90, compiled from its sources by oracle/Makefile) on two cases: its
output file (Poisson noise, and in case b binning and gamma) and the same program with POIDEV
replaced by the identity (noise-free). Here the oracle (oracle/ctem_oracle.c) is pinned to both,
and the library's host routines (Euler angles, the axis-angle rotation, POIDEV, binning, the file
readers) to the program's results; the GPU patterns are checked in test_gpu_detector.py."""
import ast
import os
import subprocess

import numpy as np
import pytest
import torch

from conftest import REPO, load_golden


def _case(z, case):
    nml = ast.literal_eval(bytes(z[f"{case}_nml"]).decode())
    return nml, str(z[f"{case}_kind"]), z[f"{case}_angles"]


def _oracle_case(z, case):
    from oracle import oracle as O
    nml, kind, ang = _case(z, case)
    det = O.ebsd_detector(nml, z[f"{case}_accum_e"], float(z["Ehistmin"]), float(z["Ebinsize"]), float(z["MCsig"]))
    q = O.eu2qu_deg(ang) if kind == "eu" else ang.reshape(-1, 4)
    q = O.apply_axisangle(q, nml["axisangle"])
    sr = O.site_sum(z[f"{case}_sr"])
    return nml, det, q, [O.ebsd_pattern(nml, det, sr, q[a]) for a in range(len(q))]


def _scale(nml, pat, binner):
    if nml["scalingmode"] == "not":
        return pat
    return binner(pat, nml["binning"], nml["scalingmode"] == "gam", nml["gammavalue"])


@pytest.mark.parametrize("case", ["a", "b"])
def test_oracle_reproduces_reference_program(case):
    """Oracle set-up, patterns, POIDEV (state carried over the patterns), binning and gamma
    against the program's noise-free and noisy output, bit for bit."""
    from oracle import oracle as O
    z = load_golden("ebsd_det")
    nml, det, q, clean = _oracle_case(z, case)
    assert det[2] == 1 and det[3] == 3
    st = np.zeros(16, np.int64)
    st[0] = -1
    for a, c in enumerate(clean):
        noisy = O.poisson_noise(c, st)
        assert np.array_equal(_scale(nml, c, O.bin_pattern).ravel(), z[f"{case}_clean"][a]), (case, a)
        assert np.array_equal(_scale(nml, noisy, O.bin_pattern).ravel(), z[f"{case}_noisy"][a]), (case, a)


def test_oracle_mistakes_are_caught():
    """Plausible slips in the restatement change the result: the plain sum instead of the
    runtime's compensated one in the normalisation shows in case a's last pattern (38 pixels),
    a swapped bilinear weight everywhere."""
    from oracle import oracle as O
    z = load_golden("ebsd_det")
    nml, det, q, clean = _oracle_case(z, "a")
    rg, wdet, emin, emax, nel, pref = det
    # swap dx and dy: transpose the master's axes
    sr = O.site_sum(z["a_sr"])
    swapped = O.ebsd_pattern(nml, det, np.ascontiguousarray(np.swapaxes(sr, 1, 2)), q[1])
    assert not np.array_equal(swapped.ravel(), z["a_clean"][1])
    # a prefactor one count of num_el off changes every pixel
    off = O.ebsd_pattern(nml, (rg, wdet, emin, emax, nel, pref * nel / (nel + 1)), sr, q[1])
    assert (off.ravel() != z["a_clean"][1]).mean() > 0.5


def test_library_host_routines_match_reference():
    """eu2qu, the axis-angle rotation, POIDEV and binning/gamma of the library (host code, no GPU)
    against the oracle and the program's output."""
    from ctemsoft_amd import detector as D
    from oracle import oracle as O
    z = load_golden("ebsd_det")
    for case in ("a", "b"):
        nml, kind, ang = _case(z, case)
        if kind == "eu":
            qd = D.eu2qu_deg(ang)
            assert np.array_equal(qd, O.eu2qu_deg(ang))
        else:
            qd = ang.reshape(-1, 4)
        qd = D.apply_axisangle(qd, nml["axisangle"])
        _, det, qo, clean = _oracle_case(z, case)
        assert np.array_equal(qd, qo)
        st = D.new_noise_state(-1)
        for a, c in enumerate(clean):
            noisy = D.poisson_noise(c, st)
            assert np.array_equal(_scale(nml, noisy, D.bin_pattern).ravel(), z[f"{case}_noisy"][a]), (case, a)
            assert np.array_equal(_scale(nml, c, D.bin_pattern).ravel(), z[f"{case}_clean"][a]), (case, a)
    # POIDEV's small-mean branch (below 20) and zeros, against the oracle
    rng = np.random.default_rng(3)
    small = (rng.random((16, 12)) * 25.0).astype(np.float32)
    small[0, :4] = 0.0
    st1, st2 = D.new_noise_state(-7), np.zeros(16, np.int64)
    st2[0] = -7
    for _ in range(3):
        assert np.array_equal(D.poisson_noise(small, st1), O.poisson_noise(small, st2))


def test_host_routines_reject_bad_input():
    from ctemsoft_amd import detector as D
    from ctemsoft_amd import ebsd
    with pytest.raises(ebsd.CtemError):  # POIDEV never returns on NaN or infinity
        D.poisson_noise(np.array([[1.0, np.nan]], np.float32), D.new_noise_state())
    with pytest.raises(ebsd.CtemError):
        D.bin_pattern(np.zeros((6, 6), np.float32), 4)


def test_readers_roundtrip(tmp_path):
    """The namelist, angle-file and MC-file readers on files written as the reference's
    programs write them."""
    from ctemsoft_amd import detector as D
    from oracle.refrun import write_mc_energy_file
    z = load_golden("ebsd_det")
    nml, kind, ang = _case(z, "b")
    lines = []
    for k, v in nml.items():
        val = f"'{v}'" if isinstance(v, str) else (", ".join(repr(float(x)) for x in v) if isinstance(v, list) else repr(v))
        lines.append(f" {k} = {val},")
    (tmp_path / "t.nml").write_text(" &EBSDdata\n" + "\n".join(lines) + "\n /\n")
    got = D.read_namelist(str(tmp_path / "t.nml"))
    for k, v in nml.items():
        assert got[k] == (list(map(float, v)) if isinstance(v, list) else v), k
    (tmp_path / "angles.txt").write_bytes(bytes(z["b_angles_txt"]))
    kind2, vals = D.read_angles(str(tmp_path / "angles.txt"))
    assert kind2 == "quats" and np.array_equal(vals, ang)
    write_mc_energy_file(str(tmp_path / "mc.data"), z["b_accum_e"], 18.0, 1.0, 70.0)
    mc = D.read_mc_energy(str(tmp_path / "mc.data"))
    assert np.array_equal(mc.accum_e, z["b_accum_e"]) and mc.nsx == 20 and mc.numEbins == 3
    assert (mc.Ehistmin, mc.Ebinsize, mc.MCsig) == (18.0, 1.0, 70.0)


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error path")
def test_detector_without_gpu_fails_cleanly():
    from ctemsoft_amd import detector as D
    from ctemsoft_amd import ebsd
    z = load_golden("ebsd_det")
    nml, _, _ = _case(z, "a")
    g = D.Geometry(*(nml[k] for k in ("L", "thetac", "delta", "numsx", "numsy", "xpc", "ypc", "energymin",
                                       "energymax", "beamcurrent", "dwelltime")))
    with pytest.raises(ebsd.CtemError) as e:
        D.Detector(g, D.MCEnergy(z["a_accum_e"], 18.0, 1.0, 70.0))
    assert e.value.code == -2


def test_lambert_math_exhaustive(tmp_path):
    """ctem_lambert.h's atanf against the host C library on every float in [-1, 1] and beyond
    (tests/native/lambert_check.cpp)."""
    exe = str(tmp_path / "lambert_check")
    src = os.path.join(REPO, "tests", "native", "lambert_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                    "-Wno-unknown-pragmas", src, "-o", exe, "-lm"], check=True)
    env = dict(os.environ, OMP_NUM_THREADS=str(min(8, os.cpu_count() or 1)))
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "atanf in [-1,1]: 0 of 2130706434 differ" in r.stdout
