"""The batched launch pattern: a plan captured into a HIP graph (ctem_plan_set_graph) and the
plans of an energy series captured into one graph (ctem_plans_run), each replayed several times
with eager runs and host synchronisation in between, every replay bit-equal to the eager run."""
import sys

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from ctemsoft_amd import ebsd
    c = ebsd.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_plan_graph_replays_equal_eager(ctx, prec):
    from ctemsoft_amd import ebsd
    fx = load_golden("fe_c1w")
    st = ebsd.Setup.from_fixture(fx)
    ctx.upload_crystal(st)
    rng = np.random.default_rng(21)
    ks = np.sort(rng.choice(len(fx["kvec"]), 384, replace=False))
    plan = ctx.plan(fx["kvec"][ks], fx["kn"][ks], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    plan.run(prec)
    ref = plan.dots().copy()
    plan.set_graph(True)
    for rep in range(3):
        plan.run(prec)
        got = plan.dots().copy()
        t = plan.timing()
        _say(f"{prec} graph replay {rep}: bit-equal {np.array_equal(got, ref)}, {t['total_ms']:.2f} ms")
        assert np.array_equal(got.view(np.int64), ref.view(np.int64))
        # eager work on the same stream between replays
        other = ctx.plan(fx["kvec"][ks[:8]], fx["kn"][ks[:8]], fx["lam"], float(fx["dz"]), float(fx["thick"]))
        other.run("fp64")
        other.close()
    plan.set_graph(False)
    plan.run(prec)
    assert np.array_equal(plan.dots().view(np.int64), ref.view(np.int64))
    plan.close()


def test_energy_series_one_graph(ctx):
    from ctemsoft_amd import ebsd, workloads
    wl = workloads.get("fe_c3w")
    arc = load_golden("fe_c3w_setups")
    fx = load_golden("fe_c1w")
    rng = np.random.default_rng(22)
    ks = np.sort(rng.choice(len(fx["kvec"]), 96, replace=False))
    plans = []
    for i in (0, 10, 20):
        view = {k[len(f"e{i}_"):]: arc[k] for k in arc.files if k.startswith(f"e{i}_")}
        st = ebsd.Setup.from_fixture(view)
        e = float(view["ekev"])
        dz = wl.dz_for(e)
        ctx.upload_crystal(st)
        kn = np.full(len(ks), 1.0 / st.mlambda)
        k = fx["kvec"][ks] * (fx["mLambda"] / st.mlambda)
        plans.append(ctx.plan(k, kn, wl.depth_weights(e), dz, float(int(wl.nz * dz))))
    for prec in ("mixed", "fp64"):
        ref = []
        for p in plans:
            p.run(prec)
            ref.append(p.dots().copy())
        for rep in range(3):
            ebsd.run_plans(plans, prec)
            for j, p in enumerate(plans):
                assert np.array_equal(p.dots().view(np.int64), ref[j].view(np.int64)), (prec, rep, j)
            _say(f"{prec}: series graph replay {rep} bit-equal over {len(plans)} bins")
            plans[1].run(prec)  # eager work between replays
            assert np.array_equal(plans[1].dots().view(np.int64), ref[1].view(np.int64))
    for p in plans:
        p.close()


def test_autotune_picks_and_caches(ctx, tmp_path, monkeypatch):
    """The launch autotuner times its candidates, applies the fastest, stores it per GPU and
    workload signature, and a second call applies the cached choice without timing. Chunk size
    changes no result; the bucket multiple changes FP64 results only at rounding level."""
    from ctemsoft_amd import autotune, ebsd
    monkeypatch.setenv("CTEM_TUNE_CACHE", str(tmp_path / "tune.json"))
    fx = load_golden("fe_c1w")
    ctx.upload_crystal(ebsd.Setup.from_fixture(fx))
    rng = np.random.default_rng(23)
    ks = np.sort(rng.choice(len(fx["kvec"]), 512, replace=False))
    plan = ctx.plan(fx["kvec"][ks], fx["kn"][ks], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    plan.run("fp64")
    ref = plan.dots().copy()
    cands = [(32, 24.0), (32, 0.25), (64, 24.0)]
    setting, times = autotune.tune(ctx, [plan], "fp64", candidates=cands, reps=1, log=_say)
    assert times is not None and set(times) == set(cands) and setting in cands
    assert (tmp_path / "tune.json").exists()
    again, t2 = autotune.tune(ctx, [plan], "fp64", candidates=cands, log=_say)
    assert t2 is None and again == setting
    plan.set_tuning(32, 0.25)  # small chunks: many launches, same results bit for bit
    plan.run("fp64")
    assert np.array_equal(plan.dots().view(np.int64), ref.view(np.int64))
    plan.set_tuning(64, 24.0)
    plan.run("fp64")
    assert np.allclose(plan.dots(), ref, rtol=1e-12, atol=0)
    plan.close()
