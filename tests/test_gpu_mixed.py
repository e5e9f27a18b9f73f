"""Mixed-precision form against the FP64 form (itself pinned to the reference within 1e-10 by
test_gpu_parity.py): per-direction intensities within 1e-4 relative."""
import sys

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

TOL = 1e-4


@pytest.fixture(scope="module")
def ctx():
    from ctemsoft_amd import ebsd
    c = ebsd.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name,nsample", [("fe_c1w", None), ("ni_c0w", None), ("fe_c1", 2048), ("ni_c0", None),
                                          ("al2o3_c2w", None), ("al2o3_c2", 2048), ("fe_c3w_10.0kV", 512),
                                          ("fe_c3w_20.0kV", 512)])
def test_mixed_matches_fp64(ctx, name, nsample):
    """All directions of the fe_c1w and al2o3_c2w fixtures (one hemisphere each), samples of the
    rest. Directions whose FP64 result is not finite (the reference's Bethe blow-up, a handful
    on corundum) must be the same in both forms and are left out of the error."""
    from ctemsoft_amd import ebsd
    fx = load_golden(name)
    # corundum: the full-box LUT (a reference-sized one does not hold every g-h of its list)
    st = ebsd.Setup.from_fixture(fx, load_golden("al2o3_c2") if name.startswith("al2o3") else None)
    ctx.upload_crystal(st)
    nk = len(fx["kvec"])
    ks = np.arange(nk) if nsample is None else np.sort(np.random.default_rng(5).choice(nk, nsample, replace=False))
    plan = ctx.plan(fx["kvec"][ks], fx["kn"][ks], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    plan.run("fp64")
    d64 = plan.dots().copy()
    plan.run("mixed")
    dmx = plan.dots().copy()
    fin = np.isfinite(d64).all(axis=-1) if d64.ndim > 1 else np.isfinite(d64)
    fin_m = np.isfinite(dmx).all(axis=-1) if dmx.ndim > 1 else np.isfinite(dmx)
    assert np.array_equal(fin, fin_m)
    rel = np.abs(dmx[fin] - d64[fin]) / np.abs(d64[fin])
    print(f"{name}: {len(ks)} directions ({int((~fin).sum())} not finite), mixed vs fp64 max rel {rel.max():.2e}, "
          f"mean {rel.mean():.2e}", flush=True)
    sys.stdout.flush()
    assert rel.max() <= TOL
    plan.close()


@pytest.mark.parametrize("taylor,square", [(3, 6), (6, 6), (3, 16), (16, 6)])
def test_split_schemes_match_fp64(ctx, taylor, square):
    """Every split scheme (3 = BF16x3, 6 = BF16x6, 16 = FP16x3) within the bar on a seeded
    sample of the Fe workload."""
    from ctemsoft_amd import ebsd
    fx = load_golden("fe_c1w")
    ctx.upload_crystal(ebsd.Setup.from_fixture(fx))
    ks = np.sort(np.random.default_rng(11).choice(len(fx["kvec"]), 512, replace=False))
    plan = ctx.plan(fx["kvec"][ks], fx["kn"][ks], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    plan.run("fp64")
    d64 = plan.dots().copy()
    plan.set_mixed_terms(taylor, square)
    plan.run("mixed")
    rel = np.abs(plan.dots() - d64) / np.abs(d64)
    print(f"schemes {taylor}/{square}: max rel {rel.max():.2e}", flush=True)
    assert rel.max() <= TOL
    plan.close()


def test_split_scheme_rejects_unknown(ctx):
    from ctemsoft_amd import ebsd
    fx = load_golden("ni_c0")
    ctx.upload_crystal(ebsd.Setup.from_fixture(fx))
    plan = ctx.plan(fx["kvec"][:4], fx["kn"][:4], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    with pytest.raises(ebsd.CtemError):
        plan.set_mixed_terms(4, 6)
    plan.close()


@pytest.mark.parametrize("name", ["fe_c1w", "al2o3_c2w"])
def test_full_products_match_fp64(ctx, name):
    """The split mixed form with symmetric products switched off (full products, the path of
    crystals without inversion symmetry) on every direction, and the symmetric products
    against it."""
    from ctemsoft_amd import ebsd
    fx = load_golden(name)
    st = ebsd.Setup.from_fixture(fx, load_golden("al2o3_c2") if name.startswith("al2o3") else None)
    ctx.upload_crystal(st)
    plan = ctx.plan(fx["kvec"], fx["kn"], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    plan.run("fp64")
    d64 = plan.dots().copy()
    assert plan.set_symmetry(1)
    plan.run("mixed")
    dsym = plan.dots().copy()
    assert not plan.set_symmetry(0)
    plan.run("mixed")
    dfull = plan.dots().copy()
    fin = np.isfinite(d64)
    assert np.array_equal(fin, np.isfinite(dfull)) and np.array_equal(fin, np.isfinite(dsym))
    rel = np.abs(dfull[fin] - d64[fin]) / np.abs(d64[fin])
    rs = np.abs(dsym[fin] - dfull[fin]) / np.abs(d64[fin])
    print(f"{name}: full products vs fp64 max rel {rel.max():.2e}; symmetric vs full {rs.max():.2e}", flush=True)
    assert rel.max() <= TOL
    assert rs.max() <= TOL
    plan.close()


def test_symmetry_detection(ctx):
    """Symmetric products are taken only for an inversion-symmetric LUT: an Fe LUT with one
    entry of a +-g pair changed by 1e-3 of the max turns them off."""
    from ctemsoft_amd import ebsd
    fx = load_golden("fe_c1w")
    st = ebsd.Setup.from_fixture(fx)
    ctx.upload_crystal(st)
    plan = ctx.plan(fx["kvec"][:4], fx["kn"][:4], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    assert plan.set_symmetry(1)
    assert not plan.set_symmetry(0)
    plan.close()
    lut = st.lut.copy()
    cell = 3 * st.lut_dim[0] + 5  # a cell off the centre of the box
    st.lut[2 * cell + 1] += 1e-3 * np.abs(lut).max()
    ctx.upload_crystal(st)
    plan = ctx.plan(fx["kvec"][:4], fx["kn"][:4], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    assert not plan.set_symmetry(1)
    plan.close()
    st.lut = lut
