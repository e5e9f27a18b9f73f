"""EBSD detector patterns on the GPU (include/ctem_detector.h) against the reference's
detector-pattern program (tests/golden/ebsd_det.npz: src_programs/CTEMEBSD.f90 compiled from its
sources, its output file and its noise-free patterns) and the oracle, bit for bit."""
import ast
import os
import struct

import numpy as np
import pytest
import torch

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _case(z, case):
    nml = ast.literal_eval(bytes(z[f"{case}_nml"]).decode())
    return nml, str(z[f"{case}_kind"]), z[f"{case}_angles"]


def _geometry(nml):
    from ctemsoft_amd import detector as D
    return D.Geometry(*(nml[k] for k in ("L", "thetac", "delta", "numsx", "numsy", "xpc", "ypc", "energymin",
                                         "energymax", "beamcurrent", "dwelltime")))


@pytest.mark.parametrize("case", ["a", "b"])
def test_patterns_equal_reference_program(case):
    """GPU patterns (noise-free) equal the oracle's; with the library's POIDEV, binning and gamma
    they equal the program's noise-free and noisy output."""
    from ctemsoft_amd import detector as D
    from oracle import oracle as O
    z = load_golden("ebsd_det")
    nml, kind, ang = _case(z, case)
    print(f"case {case}: {nml['numsx']}x{nml['numsy']}, {len(ang)} orientations", flush=True)
    det = D.Detector(_geometry(nml), D.MCEnergy(z[f"{case}_accum_e"], float(z["Ehistmin"]), float(z["Ebinsize"]),
                                                 float(z["MCsig"])))
    try:
        info = det.info()
        odet = O.ebsd_detector(nml, z[f"{case}_accum_e"], float(z["Ehistmin"]), float(z["Ebinsize"]),
                               float(z["MCsig"]))
        assert (info["Emin"], info["Emax"], info["num_el"]) == (odet[2], odet[3], odet[4])
        assert info["prefactor"] == odet[5]
        det.set_master(z[f"{case}_sr"])
        q = D.eu2qu_deg(ang) if kind == "eu" else ang.reshape(-1, 4)
        q = D.apply_axisangle(q, nml["axisangle"])
        pats = det.patterns(q)
        sr = O.site_sum(z[f"{case}_sr"])
        st = D.new_noise_state(-1)
        for a in range(len(q)):
            assert np.array_equal(pats[a], O.ebsd_pattern(nml, odet, sr, q[a])), (case, a)
            noisy = D.poisson_noise(pats[a], st)
            if nml["scalingmode"] != "not":
                c = D.bin_pattern(pats[a], nml["binning"], nml["scalingmode"] == "gam", nml["gammavalue"])
                noisy = D.bin_pattern(noisy, nml["binning"], nml["scalingmode"] == "gam", nml["gammavalue"])
            else:
                c = pats[a]
            assert np.array_equal(c.ravel(), z[f"{case}_clean"][a]), (case, a)
            assert np.array_equal(noisy.ravel(), z[f"{case}_noisy"][a]), (case, a)
    finally:
        det.close()


@pytest.mark.parametrize("case", ["a", "b"])
def test_run_program_writes_the_reference_output(tmp_path, case):
    """The drop-in runner on the program's own input files writes the program's output file:
    the same header and the same float32 patterns."""
    from ctemsoft_amd import detector as D
    from ctemsoft_amd.masterfile import write_master
    from oracle.refrun import write_mc_energy_file
    z = load_golden("ebsd_det")
    nml, kind, ang = _case(z, case)
    write_mc_energy_file(str(tmp_path / "mc.data"), z[f"{case}_accum_e"], float(z["Ehistmin"]),
                         float(z["Ebinsize"]), float(z["MCsig"]))
    sr = z[f"{case}_sr"]
    write_master(str(tmp_path / "master.data"), sr, z["ekevs"], [28] * sr.shape[0])
    (tmp_path / "angles.txt").write_bytes(bytes(z[f"{case}_angles_txt"]))

    def fmt(v):
        if isinstance(v, str):
            return f"'{v}'"
        if isinstance(v, list):
            return ", ".join(repr(float(x)) for x in v)
        return repr(v)
    (tmp_path / "CTEMEBSD.nml").write_text(" &EBSDdata\n" + "".join(f" {k} = {fmt(v)},\n" for k, v in nml.items())
                                           + " /\n")
    out = D.run_program(str(tmp_path / "CTEMEBSD.nml"))
    with open(out, "rb") as f:
        data = f.read()
    (n,) = struct.unpack_from("<i", data, 0)
    assert struct.unpack_from("<3i", data, 4) == tuple(int(v) for v in z[f"{case}_header"])
    got = D.read_output(out).reshape(len(ang), -1)
    assert np.array_equal(got, z[f"{case}_noisy"])


def test_full_size_detector_against_oracle():
    """640 x 480 pixels, 11 orientations (not a multiple of the 4 per thread), a 5-bin master of
    npx 150 with two sites, energy window of all bins: every pixel bit-equal to the oracle."""
    from ctemsoft_amd import detector as D
    from oracle import oracle as O
    rng = np.random.default_rng(11)
    npx, nE, nsx = 150, 5, 35
    yy, xx = np.mgrid[-npx:npx + 1, -npx:npx + 1]
    base = (np.cos(xx / 9.0) * np.sin(yy / 13.0) + 1.5).astype(np.float32)
    srtmp = np.stack([np.stack([base * np.float32(1 + 0.1 * e + 0.3 * s) +
                                rng.random(base.shape, dtype=np.float32) * np.float32(0.02) for e in range(nE)])
                      for s in range(2)]).astype(np.float32)
    ys, xs = np.mgrid[-nsx:nsx + 1, -nsx:nsx + 1]
    acc = np.stack([np.round(3000 * np.exp(-(xs ** 2 + ys ** 2) / 900.0) * (e + 1) + rng.integers(0, 9, xs.shape))
                    for e in range(nE)], axis=-1).astype(np.int32)
    nml = dict(L=15500.0, thetac=10.0, delta=50.0, numsx=640, numsy=480, xpc=1.5, ypc=-4.0, energymin=10.0,
               energymax=30.0, beamcurrent=14.513, dwelltime=100.0)
    mc = D.MCEnergy(acc, 16.0, 1.0, 70.0)
    det = D.Detector(_geometry(nml), mc)
    try:
        det.set_master(srtmp)
        q = rng.normal(size=(11, 4))
        q = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
        pats = det.patterns(q)
        print(f"11 patterns of 640x480 in {det.last_ms():.3f} ms", flush=True)
        odet = O.ebsd_detector(nml, acc, 16.0, 1.0, 70.0)
        assert (odet[2], odet[3]) == (1, nE)
        sr = O.site_sum(srtmp)
        for a in range(len(q)):
            ref = O.ebsd_pattern(nml, odet, sr, q[a])
            assert np.array_equal(pats[a], ref), (a, int((pats[a] != ref).sum()))
        # the device-buffer entry gives the same patterns
        dq = torch.from_numpy(q).cuda()
        dp = torch.empty((len(q), 480, 640), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        det.patterns_device(len(q), dq.data_ptr(), dp.data_ptr())
        det.sync()
        assert np.array_equal(dp.cpu().numpy(), pats)
    finally:
        det.close()


def test_detector_errors():
    from ctemsoft_amd import detector as D
    from ctemsoft_amd import ebsd
    z = load_golden("ebsd_det")
    nml, kind, ang = _case(z, "a")
    det = D.Detector(_geometry(nml), D.MCEnergy(z["a_accum_e"], 18.0, 1.0, 70.0))
    try:
        with pytest.raises(ebsd.CtemError) as e:  # no master yet
            det.patterns(np.array([[1, 0, 0, 0]], np.float32))
        assert e.value.code == -3
        with pytest.raises(ebsd.CtemError) as e:  # energy bins differ from the histograms'
            det.set_master(z["a_sr"][:, :2])
        assert e.value.code == -1
        det.set_master(z["a_sr"])
        assert det.patterns(np.zeros((0, 4), np.float32)).shape == (0, 48, 64)
        assert det.patterns(np.array([[1, 0, 0, 0]], np.float32)).shape == (1, 48, 64)
    finally:
        det.close()
