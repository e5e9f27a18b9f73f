"""GPU parity: the HIP path, called through the C-ABI, against the reference's golden
vectors (tests/golden/*.npz from the compiled reference) and the oracle.

Bars: beam sets bit-exact (indices and float32 excitation errors); DynMat to 1e-15;
FP64 Lgh within 1e-10 (relative Frobenius) of the reference's eigen-based CalcLgh3; the
float32 master-pattern values equal to the reference program's, pixel for pixel.
"""
import sys

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

FIXTURES = ["ni_c0", "ni_c0w", "fe_c1", "fe_c1w", "al2o3_c2", "al2o3_c2w",
            "fe_c3w_10.0kV", "fe_c3w_15.0kV", "fe_c3w_20.0kV"]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from ctemsoft_amd import ebsd
    c = ebsd.Context(0)
    _say("device:", c.device_info())
    yield c
    c.close()


def _plan_for(ctx, fx, ks):
    from ctemsoft_amd import ebsd
    st = ebsd.Setup.from_fixture(fx)
    ctx.upload_crystal(st)
    ks = np.asarray(ks)
    return ctx.plan(fx["kvec"][ks], fx["kn"][ks], fx["lam"], float(fx["dz"]), float(fx["thick"])), st


@pytest.mark.parametrize("name", FIXTURES)
def test_beam_sets_bit_exact(ctx, name):
    fx = load_golden(name)
    _say("beam sets", name, len(fx["beams_ik"]), "directions")
    plan, _ = _plan_for(ctx, fx, fx["beams_ik"])
    ns, nw = plan.select()
    assert np.array_equal(ns, fx["beams_nns"])
    assert np.array_equal(nw, fx["beams_nnw"])
    so = np.concatenate([[0], np.cumsum(ns)]); wo = np.concatenate([[0], np.cumsum(nw)])
    for j in range(len(ns)):
        s, ssg, w, wsg = plan.beams(j)
        assert np.array_equal(s, fx["beams_strong"][so[j]:so[j + 1]])
        assert np.array_equal(w, fx["beams_weak"][wo[j]:wo[j + 1]])
        assert np.array_equal(ssg.view(np.int32), fx["beams_strongsg"][so[j]:so[j + 1]].view(np.int32))
        assert np.array_equal(wsg.view(np.int32), fx["beams_weaksg"][wo[j]:wo[j + 1]].view(np.int32))
    plan.close()


def test_southern_hemisphere_against_reference(ctx):
    """The library's southern half of the corundum workload (al2o3_c2w; the program computes the
    northern half only) against the reference's own beam loop body on the same wave vectors
    (tests/golden/al2o3_c2w_south.npz): beam sets bit-exact on 34 directions; on the two where a
    weak beam lies ~3e-6 from the Bragg condition the reference's Lgh is not finite, and the
    library's intensities are not finite there in either form, while they are finite on every
    other direction."""
    from ctemsoft_amd import ebsd
    fx = load_golden("al2o3_c2w")
    z = load_golden("al2o3_c2w_south")
    ctx.upload_crystal(ebsd.Setup.from_fixture(fx, load_golden("al2o3_c2")))
    plan = ctx.plan(z["kvec"], z["kn"], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    ns, nw = plan.select()
    assert np.array_equal(ns, z["beams_nns"]) and np.array_equal(nw, z["beams_nnw"])
    so = np.concatenate([[0], np.cumsum(ns)]); wo = np.concatenate([[0], np.cumsum(nw)])
    for j in range(len(ns)):
        s, ssg, w, wsg = plan.beams(j)
        assert np.array_equal(s, z["beams_strong"][so[j]:so[j + 1]])
        assert np.array_equal(w, z["beams_weak"][wo[j]:wo[j + 1]])
        assert np.array_equal(ssg.view(np.int32), z["beams_strongsg"][so[j]:so[j + 1]].view(np.int32))
        assert np.array_equal(wsg.view(np.int32), z["beams_weaksg"][wo[j]:wo[j + 1]].view(np.int32))
    blow = np.isin(z["beams_ik"], z["blow_ik"])
    assert not z["blow_south_lgh_finite"].any()
    for prec in ("fp64", "mixed"):
        plan.run(prec)
        fin = np.isfinite(plan.dots()).all(axis=-1)
        _say(f"south {prec}: {int((~fin).sum())} of {len(fin)} directions not finite")
        assert np.array_equal(fin, ~blow), prec
    plan.close()


@pytest.mark.parametrize("name", FIXTURES)
def test_dynmat_and_lgh_fp64(ctx, name):
    fx = load_golden(name)
    ks = fx["dyn_ik"]
    _say("dynmat/lgh", name, list(ks))
    plan, _ = _plan_for(ctx, fx, ks)
    plan.select()
    for j in range(len(ks)):
        ref = fx[f"dyn{j}_dynmat"]
        dm = plan.dynmat(j)
        assert dm.shape == ref.shape
        err = np.abs(dm - ref).max() / np.abs(ref).max()
        assert err <= 1e-15, (name, int(ks[j]), err)
    for j in range(len(ks)):
        plan.run("fp64", keep_lgh=j)
        L = plan.lgh()
        Lr = fx[f"dyn{j}_lgh"]
        err = np.linalg.norm(L - Lr) / np.linalg.norm(Lr)
        _say(f"  k={int(ks[j])} n={L.shape[0]} lgh rel err {err:.2e}")
        assert err < 1e-10, (name, int(ks[j]), err)
    plan.close()


@pytest.mark.parametrize("name", ["ni_c0", "ni_c0w"])
def test_master_pattern_matches_reference_program(ctx, name):
    """Whole Lambert master pattern of the reference program (float32), every direction of
    the workload, scattered by the GPU."""
    from ctemsoft_amd import ebsd
    fx = load_golden(name)
    nk = len(fx["kij"])
    _say("master", name, nk, "directions")
    plan, st = _plan_for(ctx, fx, np.arange(nk))
    plan.run("fp64")
    sr_ref = fx["master_sr"][:, 0]
    npx = (sr_ref.shape[-1] - 1) // 2
    sr = plan.scatter(fx["kij"], npx, npx, int(fx["isym"]))
    diff = sr.view(np.int32).astype(np.int64) - sr_ref.view(np.int32).astype(np.int64)
    nbad = int(np.count_nonzero(diff))
    _say(f"  pixels {sr.size}, not bit-equal {nbad}, max ulp {int(np.abs(diff).max())}")
    # the GPU Lgh differs from the eigen route by ~1e-14 relative; a float32 rounding flip
    # needs the double sum within that of a rounding boundary
    assert nbad <= max(2, sr.size // 20000)
    assert np.abs(diff).max() <= 1
    plan.close()


def test_calclgh3_dropin(ctx):
    fx = load_golden("fe_c1w")
    for j, k in enumerate(fx["dyn_ik"][:2]):
        L = ctx.calclgh3(fx[f"dyn{j}_dynmat"], float(fx["thick"]), float(fx["kn"][k]), float(fx["dz"]), fx["lam"])
        Lr = fx[f"dyn{j}_lgh"]
        err = np.linalg.norm(L - Lr) / np.linalg.norm(Lr)
        _say(f"calclgh3 k={int(k)} err {err:.2e}")
        assert err < 1e-10


def test_svals_against_oracle_random_directions(ctx, oracle_mod):
    """Directions the fixtures do not hold: random wave vectors near the Fe wedge, oracle
    (extended-precision expm) as the checker."""
    fx = load_golden("fe_c1w")
    rng = np.random.default_rng(42)
    ks = rng.choice(len(fx["kvec"]), 6, replace=False)
    plan, _ = _plan_for(ctx, fx, ks)
    plan.run("fp64")
    d = plan.dots()[:, 0]
    cr = oracle_mod.Crystal(fx)
    for j, k in enumerate(ks):
        ref, _, _ = oracle_mod.master_direction(cr, fx["kvec"][k], fx["kn"][k], fx["lam"], float(fx["dz"]), float(fx["thick"]))
        assert abs(d[j] - ref[0]) <= 1e-11 * abs(ref[0]), (int(k), d[j], ref[0])
    plan.close()


def test_svals_two_sites_against_oracle(ctx, oracle_mod):
    """Corundum (R-3c, hexagonal axes): two sites (Al orbit 12, O orbit 18), one svals value per
    site, random directions of the hexagonal Lambert grid against the oracle."""
    fx = load_golden("al2o3_c2")
    rng = np.random.default_rng(7)
    ks = rng.choice(len(fx["kvec"]), 8, replace=False)
    plan, _ = _plan_for(ctx, fx, ks)
    plan.run("fp64")
    d = plan.dots()
    sv = plan.svals()
    assert d.shape == (len(ks), 2) and sv.shape == (len(ks), 2)
    cr = oracle_mod.Crystal(fx)
    nat = sum(len(fx[f"site{p}_orbit"]) for p in range(int(fx["nsite"])))
    assert nat == 30
    for j, k in enumerate(ks):
        ref, _, _ = oracle_mod.master_direction(cr, fx["kvec"][k], fx["kn"][k], fx["lam"], float(fx["dz"]), float(fx["thick"]))
        for p in range(2):
            assert abs(d[j, p] - ref[p]) <= 1e-11 * abs(ref[p]), (int(k), p, d[j, p], ref[p])
            svr = oracle_mod.svals_f32(ref[p], nat)
            # a float32 rounding flip needs the double sums within ~1e-14 of a boundary
            assert abs(int(sv[j, p].view(np.int32)) - int(svr.view(np.int32))) <= 1
    plan.close()


def test_full_box_lut_repeats_reference_inside_its_box(ctx):
    """Corundum with the LUT over the box that holds every g-h of the list (CalcUcg of the
    reference's own modules): directions whose beam pairs stay inside the reference's LUT box
    give the reference's DynMat."""
    from ctemsoft_amd import ebsd
    fx = load_golden("al2o3_c2")
    st = ebsd.Setup.from_fixture(fx, fx)
    assert st.lut_dim.tolist() == [73, 73, 105]
    ctx.upload_crystal(st)
    ks = fx["dyn_ik"]
    plan = ctx.plan(fx["kvec"][ks], fx["kn"][ks], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    plan.select()
    for j in range(len(ks)):
        ref = fx[f"dyn{j}_dynmat"]
        err = np.abs(plan.dynmat(j) - ref).max() / np.abs(ref).max()
        assert err <= 1e-15, (int(ks[j]), err)
    plan.close()


def test_south_hemisphere_mirror(ctx):
    """Fe (Im-3m) is centrosymmetric with a mirror normal to c: the southern hemisphere
    pattern equals the northern one. The mirrored wave vectors differ from the northern
    ones in the last bits (the structure matrix carries cos(90 deg) ~ 6e-17 off-diagonal
    terms), and the reference rounds them to float32 before the excitation errors, so the
    two agree to float32 input precision, not to double rounding."""
    from ctemsoft_amd import ebsd
    fx = load_golden("fe_c1")
    st = ebsd.Setup.from_fixture(fx)
    ctx.upload_crystal(st)
    ijn, kn_, knn = ebsd.kvectors(20, 20, st.isym, 0, st.dsm, st.mlambda, hemisphere=1)
    ijs, ks_, kns = ebsd.kvectors(20, 20, st.isym, 0, st.dsm, st.mlambda, hemisphere=-1)
    assert np.array_equal(ijn, ijs)
    out = []
    for k, kn in ((kn_, knn), (ks_, kns)):
        p = ctx.plan(k, kn, fx["lam"], float(fx["dz"]), float(fx["thick"]))
        p.run("fp64")
        out.append(p.dots()[:, 0])
        p.close()
    assert np.allclose(out[0], out[1], rtol=2e-6, atol=0)


@pytest.mark.parametrize("name,count", [("ni_c0", 256), ("fe_c1", 256), ("al2o3_c2", 256), ("ni_c0w", 256),
                                        ("fe_c1w", 64), ("al2o3_c2w", 48)])
def test_dots_many_directions_against_oracle(ctx, oracle_mod, name, count):
    """FP64 dots of seeded directions of each workload against the oracle (extended-precision
    Lgh), every site; the oracle runs on a thread pool (its C calls release the interpreter)."""
    from concurrent.futures import ThreadPoolExecutor
    from ctemsoft_amd import ebsd
    fx = load_golden(name)
    # corundum: the LUT over the box of every g-h of its hexagonal list, in both
    full = load_golden("al2o3_c2") if name.startswith("al2o3") else None
    st = ebsd.Setup.from_fixture(fx, full)
    ctx.upload_crystal(st)
    rng = np.random.default_rng(1000 + count)
    ks = np.sort(rng.choice(len(fx["kvec"]), min(count, len(fx["kvec"])), replace=False))
    _say("dots vs oracle", name, len(ks), "directions")
    plan = ctx.plan(fx["kvec"][ks], fx["kn"][ks], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    plan.run("fp64")
    d = plan.dots()
    plan.close()
    cr = oracle_mod.Crystal(fx, full)
    def one(k):
        return oracle_mod.master_direction(cr, fx["kvec"][k], fx["kn"][k], fx["lam"], float(fx["dz"]),
                                           float(fx["thick"]))[0]
    with ThreadPoolExecutor(16) as ex:
        ref = np.array(list(ex.map(one, ks)))
    fin = np.isfinite(ref).all(1)
    rel = np.abs(d[fin] - ref[fin]) / np.abs(ref[fin])
    _say(f"  max rel {rel.max():.2e}, non-finite in the oracle {int((~fin).sum())}")
    assert rel.max() <= 1e-11
    assert np.array_equal(np.isfinite(d), np.isfinite(ref))


@pytest.mark.parametrize("name", ["fe_c1w", "fe_c3w_10.0kV"])
def test_fp64_dots_against_reference_on_many_directions(ctx, name):
    """FP64 intensities against the reference's own CalcLgh3 (eigen-based) on 128 seeded fe_c1w
    directions and 64 at 10 kV (tests/golden/ref_dots.npz: the reference's Lgh contracted with
    S_p), within 1e-10 relative."""
    fx = load_golden(name)
    z = load_golden("ref_dots")
    ks, ref = z[f"{name}_ik"], z[f"{name}_dots"]
    plan, _ = _plan_for(ctx, fx, ks)
    plan.run("fp64")
    d = plan.dots().reshape(ref.shape)
    plan.close()
    rel = np.abs(d - ref) / np.abs(ref)
    _say(f"{name}: FP64 vs reference dots on {len(ks)} directions, max rel {rel.max():.2e}")
    assert rel.max() <= 1e-10


def test_hexagonal_scatter_keeps_last_direction(ctx):
    """Corundum (Laue group 12 on the hexagonal grid, where directions share pixels): the GPU
    scatter equals the reference's serial loop -- every direction in list order overwriting its
    Laue-equivalent pixels (CTEMEBSDmaster.f90:462-479) -- bit for bit, on every pixel."""
    from ctemsoft_amd import ebsd
    fx = load_golden("al2o3_c2")
    nk = len(fx["kij"])
    st = ebsd.Setup.from_fixture(fx, fx)
    ctx.upload_crystal(st)
    plan = ctx.plan(fx["kvec"], fx["kn"], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    plan.run("fp64")
    sv = plan.svals()
    npx, npy = ebsd.lambert_dims(200, 1)
    isym = int(fx["isym"])
    assert isym == 12
    sr = plan.scatter(fx["kij"], npx, npy, isym)
    plan.close()
    ref = np.zeros_like(sr)
    shared = np.zeros(sr.shape[1:], np.int32)
    for ik in range(nk):
        q = ebsd.laue_equiv(int(fx["kij"][ik, 0]), int(fx["kij"][ik, 1]), isym)
        q = q[(np.abs(q[:, 0]) <= npx) & (np.abs(q[:, 1]) <= npy)]
        ref[:, q[:, 1] + npy, q[:, 0] + npx] = sv[ik][:, None]
        shared[q[:, 1] + npy, q[:, 0] + npx] += 1
    _say(f"  {nk} directions, pixels reached by more than one direction: {int((shared > 1).sum())}")
    assert (shared > 1).sum() > 1000
    assert np.array_equal(sr.view(np.int32), ref.view(np.int32))
