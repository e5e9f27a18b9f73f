"""Monte Carlo on the GPU (include/ctem_mc.h) against the reference program's output
(tests/golden/mc.npz: src_programs/CTEMMC.f90 compiled from its sources) and the oracle.

Both modes evaluate the program's double-precision steps in its order, with the device's log, sin,
cos, acos and atan where the program calls the C library's: a trajectory can only part from the
program's where a last-bit difference meets a threshold (the 1 keV cut-off, the surface, a bin
edge) -- none does on these runs, so the histograms are compared count for count."""
import ast
import os
import tempfile

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

RUNS = ["m1", "m2", "m3"]


def _case(z, run):
    return ast.literal_eval(bytes(z[f"{run}_xtal"]).decode()), ast.literal_eval(bytes(z[f"{run}_nml"]).decode())


def _ref(z, run):
    from oracle.refrun import read_mc_file
    with tempfile.NamedTemporaryFile(suffix=".data", delete=False) as f:
        f.write(z[f"{run}_file"].tobytes())
    try:
        return read_mc_file(f.name)
    finally:
        os.unlink(f.name)


@pytest.mark.parametrize("run", RUNS)
def test_sequential_equals_the_program(run):
    from ctemsoft_amd import mc
    z = load_golden("mc")
    x, nml = _case(z, run)
    print(f"run {run}: {nml}", flush=True)
    r = mc.simulate(x, mode="sequential", **{k: v for k, v in nml.items() if k != "nthreads"})
    ref = _ref(z, run)
    print(f"  {r['accum_e'].sum()} hits, {r['ms']:.1f} ms", flush=True)
    assert r["dropped"] == 0
    assert np.array_equal(r["accum_e"], ref["accum_e"])
    assert np.array_equal(r["accum_z"], ref["accum_z"])


@pytest.mark.parametrize("run", ["m1", "m2"])
def test_blocked_equals_the_oracle(run, oracle_mod):
    """the blocked streams, electrons in parallel, against the oracle's blocked run (electrons in
    order on the CPU), count for count; split into two electron ranges run in reverse order, the
    same histograms"""
    from ctemsoft_amd import mc, xtal
    O = oracle_mod
    z = load_golden("mc")
    x, nml = _case(z, run)
    dens, avz, ava = xtal.density(x)
    p = dict(nml, Ze=float(avz), density=float(dens), at_wt=float(ava))
    ae_o, az_o, dr_o = O.mc_run(p, mode=1)
    m = mc.MonteCarlo(p["sig"], p["omega"], p["numsx"], p["primeseed"], p["num_el"], p["EkeV"], p["Ehistmin"],
                      p["Ebinsize"], p["depthmax"], p["depthstep"], p["Ze"], p["density"], p["at_wt"])
    try:
        m.run("blocked")
        ae, az, dr = m.histograms()
        print(f"run {run}: {ae.sum()} hits in {m.last_ms():.2f} ms", flush=True)
        assert dr == dr_o
        assert np.array_equal(ae, ae_o) and np.array_equal(az, az_o)
        m.reset()
        half = p["num_el"] // 2
        m.run("blocked", first=half, count=p["num_el"] - half)
        m.run("blocked", first=0, count=half)
        ae2, az2, _ = m.histograms()
        assert np.array_equal(ae2, ae) and np.array_equal(az2, az)
    finally:
        m.close()


def test_run_program_writes_the_reference_file(tmp_path):
    """the drop-in runner on the program's namelist and crystal file writes the program's output
    file byte for byte"""
    from ctemsoft_amd import mc
    from oracle.refrun import write_xtal
    z = load_golden("mc")
    x, nml = _case(z, "m2")
    write_xtal(str(tmp_path / "crystal.xtal"), x)
    body = dict(nml, xtalname="crystal.xtal", dataname="mc.data")

    def fmt(v):
        return f"'{v}'" if isinstance(v, str) else repr(v)
    (tmp_path / "CTEMMC.nml").write_text(" &MCdata\n" + "".join(f" {k} = {fmt(v)},\n" for k, v in body.items()) + " /\n")
    out = mc.run_program(str(tmp_path / "CTEMMC.nml"))
    assert open(out, "rb").read() == z["m2_file"].tobytes()


def test_template_size_symmetry():
    """the program's template size (numsx 1501, 30 kV, 51 energy bins, 101 depth bins) with a million
    Ni electrons: the exit-direction map is mirror-symmetric across the tilt plane within Poisson
    noise (chi-square over 25 x 25-pixel blocks), every electron is accounted for, and the energy
    spectrum falls off from the beam energy"""
    from ctemsoft_amd import mc, workloads
    r = mc.simulate(workloads.get("ni_c0").xtal, mode="blocked", sig=70.0, omega=0.0, numsx=1501, primeseed=932117,
                    num_el=1_000_000, EkeV=30.0, Ehistmin=5.0, Ebinsize=0.5, depthmax=100.0, depthstep=1.0)
    ae = r["accum_e"]
    n = int(ae.sum())
    print(f"{n} hits of 1e6 electrons in {r['ms']:.1f} ms", flush=True)
    assert r["dropped"] == 0 and 200_000 < n < 800_000
    # the map is [j][i] (the program's accum_e(iE, i, j)); y -> -y reflects i: the half-disk i < 0
    # against the half-disk i > 0 read backwards, in 25 x 25-pixel blocks with counts
    m = ae.sum(axis=2).astype(float)
    blocks = lambda h: h[:1500].reshape(60, 25, 30, 25).sum(axis=(1, 3))
    a, b = blocks(m[:, :750]), blocks(m[:, 751:][:, ::-1])
    mask = (a + b) > 50
    chi = np.sum((a[mask] - b[mask]) ** 2 / (a[mask] + b[mask]))
    dof = int(mask.sum())
    assert chi < dof + 6 * np.sqrt(2 * dof), (chi, dof)
    spec = ae.sum(axis=(0, 1))
    assert spec[-1] > 0 and spec.argmax() >= len(spec) // 2


def test_mc_errors():
    from ctemsoft_amd import ebsd, mc
    with pytest.raises(ebsd.CtemError) as e:  # numsx must be odd
        mc.MonteCarlo(70.0, 0.0, 100, 1, 10, 20.0, 10.0, 1.0, 100.0, 1.0, 28.0, 8.9, 58.7)
    assert e.value.code == -1
    m = mc.MonteCarlo(70.0, 0.0, 101, 1, 10, 20.0, 10.0, 1.0, 100.0, 1.0, 28.0, 8.9, 58.7)
    try:
        with pytest.raises(ebsd.CtemError) as e:  # the single stream starts at electron 0
            m.run("sequential", first=5, count=5)
        assert e.value.code == -1
        m.run("blocked", first=0, count=0)
        ae, az, dr = m.histograms()
        assert ae.sum() == 0 and dr == 0
    finally:
        m.close()
