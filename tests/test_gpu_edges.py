"""Edge cases of the plan API on the GPU: an empty direction list, results that do not depend on
the batch a direction is run in, the shortest depth profiles (one and two depth steps, which
leave the depth recursions without a step or with a single one), and argument errors through
the C-ABI (CTEMEBSDmaster.f90:406-484 runs every direction independently, with izzmax >= 1)."""
import ctypes as C
import sys

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

E_ARG, E_STATE = -1, -3


@pytest.fixture(scope="module")
def ctx():
    from ctemsoft_amd import ebsd
    c = ebsd.Context(0)
    yield c
    c.close()


def _upload(ctx, name):
    from ctemsoft_amd import ebsd
    fx = load_golden(name)
    full = load_golden("al2o3_c2") if name.startswith("al2o3") else None
    ctx.upload_crystal(ebsd.Setup.from_fixture(fx, full))
    return fx


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_empty_direction_list(ctx, precision):
    fx = _upload(ctx, "fe_c1w")
    plan = ctx.plan(np.zeros((0, 3)), np.zeros(0), fx["lam"], float(fx["dz"]), float(fx["thick"]))
    ns, nw = plan.select()
    assert len(ns) == 0 and len(nw) == 0
    plan.run(precision)
    assert plan.dots().shape == (0, 1) and plan.svals().shape == (0, 1)
    plan.close()


@pytest.mark.parametrize("name", ["fe_c1w", "al2o3_c2w"])
def test_results_independent_of_batch(ctx, name):
    """A direction's intensities are the same bits whether it runs in a batch of 384, in the
    reversed batch or alone: every matrix of the exponential and of the depth recursion has its
    own exponents, norms and squaring count, in both forms."""
    fx = _upload(ctx, name)
    ks = np.sort(np.random.default_rng(3).choice(len(fx["kvec"]), 384, replace=False))
    args = (fx["lam"], float(fx["dz"]), float(fx["thick"]))
    for precision in ("fp64", "mixed"):
        plan = ctx.plan(fx["kvec"][ks], fx["kn"][ks], *args)
        plan.run(precision)
        d = plan.dots().copy()
        plan.close()
        rev = ks[::-1]
        plan = ctx.plan(fx["kvec"][rev], fx["kn"][rev], *args)
        plan.run(precision)
        dr = plan.dots()[::-1].copy()
        plan.close()
        assert np.array_equal(d, dr, equal_nan=True), (name, precision)
        for j in (0, 191, 383):
            plan = ctx.plan(fx["kvec"][ks[j:j + 1]], fx["kn"][ks[j:j + 1]], *args)
            plan.run(precision)
            assert np.array_equal(plan.dots()[0], d[j], equal_nan=True), (name, precision, j)
            plan.close()
        print(f"{name} {precision}: batch, reversed batch and single runs bit-equal", flush=True)
    sys.stdout.flush()


def test_one_depth_step_against_oracle(ctx, oracle_mod):
    """izzmax = 1: Lgh is the top layer's outer product only; FP64 against the oracle."""
    fx = _upload(ctx, "fe_c1w")
    ks = np.array([100, 20000, 31000])
    lam1 = fx["lam"][:1]
    plan = ctx.plan(fx["kvec"][ks], fx["kn"][ks], lam1, float(fx["dz"]), float(fx["thick"]))
    plan.run("fp64")
    d = plan.dots()[:, 0].copy()
    plan.run("mixed")
    dm = plan.dots()[:, 0].copy()
    plan.close()
    cr = oracle_mod.Crystal(fx)
    for j, k in enumerate(ks):
        ref, _, _ = oracle_mod.master_direction(cr, fx["kvec"][k], fx["kn"][k], lam1, float(fx["dz"]), float(fx["thick"]))
        assert abs(d[j] - ref[0]) <= 1e-11 * abs(ref[0]), (int(k), d[j], ref[0])
    assert np.max(np.abs(dm - d) / np.abs(d)) <= 1e-4


@pytest.mark.parametrize("name", ["fe_c1w", "al2o3_c2w"])
@pytest.mark.parametrize("nz", [1, 2, 3, 5, 6])
def test_short_depth_profiles_mixed_vs_fp64(ctx, name, nz):
    """One to six depth steps: the on-chip recursion (Fe) and the four-depth recursion
    (corundum, whose first pass gives depth 1, its E^2 pass depths 2 and 3, and whose E^4 passes
    start at depth 4) against FP64."""
    fx = _upload(ctx, name)
    ks = np.sort(np.random.default_rng(nz).choice(len(fx["kvec"]), 256, replace=False))
    plan = ctx.plan(fx["kvec"][ks], fx["kn"][ks], fx["lam"][:nz], float(fx["dz"]), float(fx["thick"]))
    plan.run("fp64")
    d64 = plan.dots().copy()
    plan.run("mixed")
    dmx = plan.dots().copy()
    plan.close()
    fin = np.isfinite(d64).all(axis=-1)
    assert np.array_equal(fin, np.isfinite(dmx).all(axis=-1))
    rel = np.abs(dmx[fin] - d64[fin]) / np.abs(d64[fin])
    print(f"{name} nz={nz}: mixed vs fp64 max rel {rel.max():.2e}", flush=True)
    assert rel.max() <= 1e-4


def test_argument_errors(ctx):
    """Bad arguments come back as status codes with a message, never as a crash."""
    from ctemsoft_amd import ebsd
    L = ebsd.lib()
    fx = _upload(ctx, "fe_c1w")
    k = np.ascontiguousarray(fx["kvec"][:4], np.float64)
    kn = np.ascontiguousarray(fx["kn"][:4], np.float64)
    lam = np.ascontiguousarray(fx["lam"], np.float32)
    dz, thick = float(fx["dz"]), float(fx["thick"])
    P = lambda a: a.ctypes.data_as(C.c_void_p)

    def create(numk, kk, nz, dz_, thick_):
        h = C.c_void_p()
        rc = L.ctem_plan_create(ctx._h, numk, P(kk), P(kn), nz, P(lam), dz_, thick_, C.byref(h))
        return rc, h

    assert create(4, k, 0, dz, thick)[0] == E_ARG            # no depth weights
    assert create(4, k, 5000, dz, thick)[0] == E_ARG         # more than 4096
    assert create(4, k, len(lam), 0.0, thick)[0] == E_ARG    # dz must be positive
    assert create(4, k, len(lam), dz, -1.0)[0] == E_ARG      # thick must be positive
    assert create(-1, k, len(lam), dz, thick)[0] == E_ARG    # negative count
    knan = k.copy()
    knan[2, 1] = np.nan
    rc, h = create(4, knan, len(lam), dz, thick)
    assert rc == E_ARG and not h.value
    assert b"non-finite" in L.ctem_last_error()

    rc, h = create(4, k, len(lam), dz, thick)
    assert rc == 0 and h.value
    dots = np.zeros(4)
    assert L.ctem_plan_get_dots(h, P(dots)) == E_STATE       # not run yet
    assert L.ctem_plan_run(h, 7, -1) == E_ARG                # unknown precision
    assert L.ctem_plan_run(h, 0, 4) == E_ARG                 # keep_lgh_ik past the list
    assert L.ctem_plan_set_symmetry(h, 2, None) == E_ARG
    assert L.ctem_plan_run(h, 0, -1) == 0                    # the plan is still usable
    assert L.ctem_plan_get_dots(h, P(dots)) == 0 and np.isfinite(dots).all()
    assert L.ctem_plan_destroy(h) == 0
    assert L.ctem_plan_run(None, 0, -1) == E_ARG
    # a crystal whose site arrays are missing, or whose orbit is empty, is refused at upload
    st = ctx.setup.struct()
    st.site_z = None
    assert L.ctem_upload_crystal(ctx._h, C.byref(st)) == E_ARG
    assert b"null site arrays" in L.ctem_last_error()
    st = ctx.setup.struct()
    norb = np.zeros(max(1, st.nsite), np.int32)
    st.site_norbit = P(norb)
    assert L.ctem_upload_crystal(ctx._h, C.byref(st)) == E_ARG


def test_site_tables_follow_the_crystal():
    """A context that alternates between crystals (Fe, two-site corundum, Fe again) gives each plan
    the structure-matrix table of its own sites: the dots equal those of a fresh context per
    crystal, bit for bit (the table is shared only between uploads with equal sites and box)."""
    from ctemsoft_amd import ebsd
    names = ["fe_c1w", "al2o3_c2", "fe_c1w"]
    shared = ebsd.Context(0)
    plans, fxs = [], []
    for name in names:
        fx = _upload(shared, name)
        n = min(16, len(fx["kn"]))
        plans.append(shared.plan(fx["kvec"][:n], fx["kn"][:n], fx["lam"], float(fx["dz"]), float(fx["thick"])))
        fxs.append(fx)
    got = []
    for p in plans:
        p.run("fp64")
        got.append(p.dots().copy())
        p.close()
    shared.close()
    for name, fx, d in zip(names, fxs, got):
        own = ebsd.Context(0)
        _upload(own, name)
        n = min(16, len(fx["kn"]))
        p = own.plan(fx["kvec"][:n], fx["kn"][:n], fx["lam"], float(fx["dz"]), float(fx["thick"]))
        p.run("fp64")
        ref = p.dots().copy()
        p.close()
        own.close()
        print(f"{name}: {n} directions, shared-context dots equal: {np.array_equal(d, ref)}", flush=True)
        assert d.shape == ref.shape and np.array_equal(d, ref)
    assert np.array_equal(got[0], got[2])
