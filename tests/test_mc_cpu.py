"""Monte Carlo on the CPU: the oracle's restatement of the trajectory loop and of the program's
random-number generator against the reference program's output (tests/golden/mc.npz: the compiled
CTEMMC and its set-up dumped by oracle/ref_mc.f90), the native CalcDensity, the generator's jump,
the output-file writer, and the blocked random streams' statistics against the program's."""
import ast
import os
import tempfile

import numpy as np
import pytest

from conftest import load_golden

RUNS = ["m1", "m2", "m3"]


def _case(z, run):
    x = ast.literal_eval(bytes(z[f"{run}_xtal"]).decode())
    nml = ast.literal_eval(bytes(z[f"{run}_nml"]).decode())
    return x, nml


def _ref(z, run):
    from oracle.refrun import read_mc_file
    with tempfile.NamedTemporaryFile(suffix=".data", delete=False) as f:
        f.write(z[f"{run}_file"].tobytes())
    try:
        return read_mc_file(f.name)
    finally:
        os.unlink(f.name)


def _params(z, run):
    x, nml = _case(z, run)
    dens, avz, ava = (float(v) for v in z[f"{run}_density"])
    return dict(nml, Ze=avz, density=dens, at_wt=ava)


@pytest.mark.parametrize("run", RUNS)
def test_density_and_generator(run, oracle_mod):
    """CalcDensity (native, float32) and the program's random stream (oracle) bit for bit; the
    generator's jump by n draws equals n draws"""
    from ctemsoft_amd import xtal
    O = oracle_mod
    z = load_golden("mc")
    x, nml = _case(z, run)
    assert np.array_equal(np.array(xtal.density(x), np.float32), z[f"{run}_density"])
    u, _ = O.mc_rng_stream(nml["primeseed"] + 1, 2000)
    assert np.array_equal(u, z[f"{run}_u"])
    for n in (1, 7, 999, 1500):
        uj, _ = O.mc_rng_stream(nml["primeseed"] + 1, 2000 - n, jump=n)
        assert np.array_equal(uj, u[n:]), n
    # a long jump against the same number of single draws
    _, s_seq = O.mc_rng_stream(nml["primeseed"] + 1, 40000)
    _, s_jmp = O.mc_rng_stream(nml["primeseed"] + 1, 0, jump=40000)
    assert s_seq == s_jmp


@pytest.mark.parametrize("run", RUNS)
def test_oracle_reproduces_the_program(run, oracle_mod):
    """single_run restated (one stream, the program's order): accum_e and accum_z equal the program's
    output file, count for count"""
    O = oracle_mod
    z = load_golden("mc")
    ref = _ref(z, run)
    ae, az, dropped = O.mc_run(_params(z, run), mode=0)
    assert dropped == 0
    assert np.array_equal(ae, ref["accum_e"]) and np.array_equal(az, ref["accum_z"])


def test_output_writer_reproduces_the_file():
    """The runner's writer on the program's own histograms gives the program's output file, byte for
    byte (records, padding, the electron count of its single loop pass)"""
    from ctemsoft_amd import mc
    z = load_golden("mc")
    for run in RUNS:
        x, nml = _case(z, run)
        ref = _ref(z, run)
        full = dict(mc.NML_DEFAULTS)
        full.update(nml, xtalname="crystal.xtal")
        with tempfile.TemporaryDirectory() as w:
            path = os.path.join(w, "out.data")
            mc.write_output(path, full, ref["accum_e"], ref["accum_z"])
            with open(path, "rb") as f:
                assert f.read() == z[f"{run}_file"].tobytes(), run


def test_read_xtal_round_trip():
    from ctemsoft_amd import xtal
    from oracle.refrun import write_xtal
    z = load_golden("mc")
    x, _ = _case(z, "m3")
    with tempfile.TemporaryDirectory() as w:
        write_xtal(os.path.join(w, "c.xtal"), x)
        y = xtal.read_xtal(os.path.join(w, "c.xtal"))
    assert y["sg"] == x["sg"] and y["system"] == x["system"] and len(y["sites"]) == len(x["sites"])
    for a, b in zip(x["sites"], y["sites"]):
        assert a["Z"] == b["Z"] and np.float32(a["occ"]) == np.float32(b["occ"])
        assert np.allclose(a["xyz"], b["xyz"]) and np.float32(a["B"]) == np.float32(b["B"])


def test_blocked_streams_match_the_program_statistically(oracle_mod):
    """Blocked streams (electron e on draws 16384 e ..) estimate the same histograms as the program's
    single stream: over 40,000 Ni electrons the backscatter count agrees within 4 standard errors
    and the energy spectrum and the exit-direction map (4 x 4 blocks) pass a chi-square test against
    the program's own run of the same size (both samples independent)"""
    O = oracle_mod
    z = load_golden("mc")
    p = _params(z, "m1")
    p["num_el"] = 40000
    ae_b, az_b, _ = O.mc_run(p, mode=1)
    ae_s, az_s, _ = O.mc_run(p, mode=0)
    n = p["num_el"]
    nb, ns = int(ae_b.sum()), int(ae_s.sum())
    q = (nb + ns) / (2 * n)
    assert abs(nb - ns) <= 4 * np.sqrt(2 * n * q * (1 - q)), (nb, ns)

    def chi2(a, b):
        a = a.astype(float).ravel()
        b = b.astype(float).ravel()
        m = (a + b) > 20
        k = a[m].sum() / b[m].sum()
        c = np.sum((a[m] - k * b[m]) ** 2 / (a[m] + k * k * b[m]))
        return c, int(m.sum())
    c, dof = chi2(ae_b.sum(axis=(0, 1)), ae_s.sum(axis=(0, 1)))
    assert c < dof + 5 * np.sqrt(2 * dof), (c, dof)
    coarse = lambda a: a.sum(axis=2)[:100, :100].reshape(25, 4, 25, 4).sum(axis=(1, 3))
    c, dof = chi2(coarse(ae_b), coarse(ae_s))
    assert c < dof + 5 * np.sqrt(2 * dof), (c, dof)
