"""The look-up-table box of the per-energy setup built on the GPU (ctem_setup_create_gpu) against
the host's bit-exact restatement of the reference (itself pinned by tests/test_setup_native.py),
and the path run from the GPU-built setup against the reference's beam sets and the host-built
setup's intensities."""
import sys

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [("ni_c0", None), ("fe_c1w", None), ("al2o3_c2", None), ("fe_c3w", 10.0), ("fe_c3w", 20.0)]


@pytest.mark.parametrize("name,ekev", CASES)
def test_gpu_lut_against_host(name, ekev):
    from ctemsoft_amd import workloads, xtal
    wl = workloads.get(name)
    e = wl.ekev if ekev is None else ekev
    cut = (wl.weakcutoff, wl.cutoff, wl.sgcutoff)
    sh, vh = xtal.native_setup(wl.xtal, e, wl.dmin, cut)
    sg, vg = xtal.native_setup(wl.xtal, e, wl.dmin, cut, device=0)
    lh = sh.lut.view(np.complex128)
    lg = sg.lut.view(np.complex128)
    umax = np.abs(lh).max()
    dev = np.abs(lg - lh) / umax
    bits = np.mean((lg.view(np.float64) == lh.view(np.float64)))
    _say(f"{name} {e:g} kV: {lh.size} cells, kernel {vg['lut_ms']:.3f} ms, bit-equal parts {bits:.4f}, "
         f"max |dU|/max|U| {dev.max():.2e}")
    # list order, 000, wavelength: the host's
    assert np.array_equal(sg.hkl, sh.hkl)
    assert vg["mLambda"] == vh["mLambda"] and vg["DynUpz"] == vh["DynUpz"]
    # a few float32 rounding steps of the device's expf/logf/powf/sincosf, relative to max|U|
    assert dev.max() <= 2e-6
    rel = np.abs(lg - lh) / np.maximum(np.abs(lh), 1e-30)
    big = np.abs(lh) > 1e-4 * umax
    assert rel[big].max() <= 2e-5
    # double-diffraction flags of the list entries agree (no |U| lies near the 1e-10 threshold)
    assert np.array_equal(sg.dbdiff, sh.dbdiff)


@pytest.mark.parametrize("name", ["fe_c1w", "al2o3_c2"])
def test_path_from_gpu_setup(name):
    """Beam sets from the GPU-built setup equal the reference's on the fixture directions, and
    the FP64 intensities agree with those from the host-built (reference-equal) setup."""
    from ctemsoft_amd import ebsd, workloads, xtal
    fx = load_golden(name)
    wl = workloads.get(name)
    cut = (wl.weakcutoff, wl.cutoff, wl.sgcutoff)
    sg, _ = xtal.native_setup(wl.xtal, wl.ekev, wl.dmin, cut, device=0)
    sh, _ = xtal.native_setup(wl.xtal, wl.ekev, wl.dmin, cut)
    ks = fx["beams_ik"]
    ctx = ebsd.Context(0)
    out = []
    for st in (sg, sh):
        ctx.upload_crystal(st)
        plan = ctx.plan(fx["kvec"][ks], fx["kn"][ks], fx["lam"], float(fx["dz"]), float(fx["thick"]))
        ns, nw = plan.select()
        if st is sg:
            assert np.array_equal(ns, fx["beams_nns"]) and np.array_equal(nw, fx["beams_nnw"])
            so = np.concatenate([[0], np.cumsum(ns)])
            for j in range(len(ns)):
                s, ssg, w, wsg = plan.beams(j)
                assert np.array_equal(s, fx["beams_strong"][so[j]:so[j + 1]])
        if name.startswith("al2o3"):
            plan.close()
            continue  # corundum's reference-sized LUT cannot hold every g-h of its list
        plan.run("fp64")
        out.append(plan.dots().copy())
        plan.close()
    ctx.close()
    if out:
        rel = np.abs(out[0] - out[1]) / np.abs(out[1])
        _say(f"{name}: dots from the GPU-built setup vs host-built, max rel {rel.max():.2e}")
        assert rel.max() <= 1e-5
