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


@pytest.mark.parametrize("name,nsample", [("fe_c1w", 1024), ("ni_c0w", None), ("fe_c1", 2048), ("ni_c0", None)])
def test_mixed_matches_fp64(ctx, name, nsample):
    from ctemsoft_amd import ebsd
    fx = load_golden(name)
    st = ebsd.Setup.from_fixture(fx)
    ctx.upload_crystal(st)
    nk = len(fx["kvec"])
    ks = np.arange(nk) if nsample is None else np.sort(np.random.default_rng(5).choice(nk, nsample, replace=False))
    plan = ctx.plan(fx["kvec"][ks], fx["kn"][ks], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    plan.run("fp64")
    d64 = plan.dots().copy()
    plan.run("mixed")
    dmx = plan.dots().copy()
    rel = np.abs(dmx - d64) / np.abs(d64)
    print(f"{name}: {len(ks)} directions, mixed vs fp64 max rel {rel.max():.2e}, mean {rel.mean():.2e}", flush=True)
    sys.stdout.flush()
    assert rel.max() <= TOL
    plan.close()
