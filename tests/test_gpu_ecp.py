"""Channeling patterns on the GPU (include/ctem_ecp.h) against the reference program's output
(tests/golden/ecp.npz: src_programs/CTEMECP.f90 compiled from its sources) and the oracle.

Tolerances: ScatMat evaluates the program's own Pade approximant (math.f90:187-315), so the float32
intensities may differ from the program's only where the double sums round to neighbouring
floats: at most 1 ulp. Blochwv integrates exactly (block-matrix exponential) where the program
rounds the exponents of its Bloch-wave integral to float32 (the default-kind cmplx() of
CTEMECP.f90:514): measured 1 ulp (1.2e-7 relative) between the exact integral and the program on
these cases; the test allows 2 ulp against the program and 1 ulp against the oracle's exact
integral."""
import numpy as np
import pytest

from ecp_cases import fixture, native_setup, oracle_crystal, run_params, setup_params, thicknesses

pytestmark = pytest.mark.gpu

RUNS = ["a", "b", "c", "d", "e", "f"]


def _ulps(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


def _plan(ctx, z, key, rp, geo):
    from ctemsoft_amd import ecp
    sp = setup_params(z, key)
    occ = [s["occ"] for s in sp["xtal"]["sites"]]
    return ecp.EcpPlan(ctx, geo["k"], geo["kn"], rp["compmode"], rp["startthick"], rp["thickinc"],
                       rp["numthick"], rp["zintstep"], geo["dynfn"], occ)


@pytest.mark.parametrize("run", RUNS)
def test_ecp_matches_the_program(run):
    from ctemsoft_amd import ebsd
    z = fixture()
    rp = run_params(z, run)
    key = rp["setup"]
    st, vals, geo = native_setup(z, key)
    ref = np.asarray(z[f"{run}_sr"])
    print(f"run {run}: {rp}, {len(geo['kn'])} directions", flush=True)
    ctx = ebsd.Context(0)
    try:
        ctx.upload_crystal(st)
        p = _plan(ctx, z, key, rp, geo)
        p.select()
        # the first direction's beams and DynMat are the program's (ECPmatrix.data)
        s0, sg0, _, _ = p.beams(0)
        assert np.array_equal(st.hkl[s0], z[f"{key}_dyn0_hkl"])
        assert np.array_equal(sg0, z[f"{key}_dyn0_sg"])
        # (the weak-beam sums run on FP64 MFMA in another order than the reference's loop)
        ref0 = np.asarray(z[f"{key}_dyn0_dynmat"])
        assert np.abs(p.dynmat(0) - ref0).max() <= 1e-15 * np.abs(ref0).max()
        p.run()
        sr = p.sr()
        u = _ulps(sr, ref)
        print(f"  max ulp {u.max()}, {int((u > 0).sum())} of {u.size} values differ", flush=True)
        assert u.max() <= (1 if rp["compmode"] == "ScatMat" else 2)
        img = p.image(geo["ij"], setup_params(z, key)["npix"])
        for ik, (i, j) in enumerate(geo["ij"]):
            npx = setup_params(z, key)["npix"]
            assert np.array_equal(img[:, j + npx, i + npx], sr[:, ik])
        if run == "e":
            assert not sr.any()
    finally:
        ctx.close()


@pytest.mark.parametrize("run", ["a", "d", "f"])
def test_blochwv_against_the_exact_integral(run, oracle_mod):
    """Blochwv against the oracle's Bloch-wave integral with double exponents: within 1 ulp"""
    from ctemsoft_amd import ebsd
    O = oracle_mod
    z = fixture()
    rp = run_params(z, run)
    key = rp["setup"]
    st, vals, geo = native_setup(z, key)
    cr = oracle_crystal(z, key)
    thick = thicknesses(rp)
    ctx = ebsd.Context(0)
    try:
        ctx.upload_crystal(st)
        p = _plan(ctx, z, key, rp, geo)
        p.run()
        sr = p.sr()
    finally:
        ctx.close()
    for ik in range(len(geo["kn"])):
        want, *_ = O.ecp_direction(cr, geo["k"][ik], geo["kn"][ik], geo["dynfn"], "Blochwv", thick, exact=True)
        assert _ulps(sr[:, ik], want).max() <= 1, (ik, sr[:, ik], want)


@pytest.mark.parametrize("mode", ["Blochwv", "ScatMat"])
def test_full_disk_symmetry(mode):
    """Ni [001] with foil normal [001], 2 npix + 1 = 65 pixels across (3,209 directions): the
    pattern has the zone's 4mm symmetry -- sr(i, j) = sr(j, i) = sr(-i, j) -- to the float32
    rounding of the directions, and every intensity is finite and positive."""
    from ctemsoft_amd import ecp, workloads
    r = ecp.compute(workloads.get("ni_c0").xtal, voltage=20000.0, k=(0, 0, 1), fn=(0, 0, 1), dmin=0.05,
                    thetac=120.0, startthick=2.0, thickinc=4.0, numthick=4, npix=32, compmode=mode,
                    zintstep=1.0, cutoffs=(4.0, 8.0, 0.05))
    sr = r["sr"]
    print(f"{mode}: {r['numk']} directions, {r['timing']['total_ms']:.2f} ms", flush=True)
    w = sr.shape[1]
    yy, xx = np.mgrid[0:w, 0:w] - 32
    disk = xx ** 2 + yy ** 2 <= 32 * 32
    assert np.all(np.isfinite(sr)) and np.all(sr[:, disk] > 0) and not sr[:, ~disk].any()
    for t in range(sr.shape[0]):
        a = sr[t]
        scale = a[disk].max()
        assert np.abs(a - a.T).max() <= 1e-4 * scale
        assert np.abs(a - a[:, ::-1]).max() <= 1e-4 * scale
        assert np.abs(a - a[::-1, :]).max() <= 1e-4 * scale


def test_ecp_errors():
    from ctemsoft_amd import ebsd, ecp
    z = fixture()
    st, vals, geo = native_setup(z, "s1")
    ctx = ebsd.Context(0)
    try:
        ctx.upload_crystal(st)
        with pytest.raises(ebsd.CtemError) as e:  # Blochwv divides by the thickness
            ecp.EcpPlan(ctx, geo["k"], geo["kn"], "Blochwv", 0.0, 2.0, 3, 1.0, geo["dynfn"], [1.0])
        assert e.value.code == -1
        with pytest.raises(ebsd.CtemError) as e:  # ScatMat needs a positive step
            ecp.EcpPlan(ctx, geo["k"], geo["kn"], "ScatMat", 2.0, 2.0, 3, 0.0, geo["dynfn"], [1.0])
        assert e.value.code == -1
        p = ecp.EcpPlan(ctx, geo["k"][:3], geo["kn"][:3], "ScatMat", 2.0, 2.0, 3, 1.0, geo["dynfn"], [1.0])
        with pytest.raises(ebsd.CtemError) as e:  # an ECP plan runs with ctem_ecp_run
            p.run_master = ebsd.Plan.run(p, "fp64")
        assert e.value.code == -3
        with pytest.raises(ebsd.CtemError) as e:  # no results before a run
            p.sr()
        assert e.value.code == -3
        p.run()
        assert p.sr().shape == (3, 3)
    finally:
        ctx.close()
