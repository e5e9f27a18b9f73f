"""The look-up-table box of the per-energy setup built on the GPU (ctem_setup_create_gpu) against
the host's restatement of the reference (itself pinned by tests/test_setup_native.py) and against
the reference's own setups, bit for bit, and the path run from the GPU-built setup against the
reference's beam sets and the host-built setup's intensities. The kernel evaluates the C
library's expf, logf, sincosf and cabsf through their restatement (ctemsoft_amd/csrc/ctem_libm.h,
pinned exhaustively by tests/test_sg_math.py), so every float32 Ucg is the reference's."""
import sys
import time

import numpy as np
import pytest

from conftest import load_golden
from test_setup_native import FIXTURES, _compare

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
    t0 = time.perf_counter()
    sh, vh = xtal.native_setup(wl.xtal, e, wl.dmin, cut)
    t1 = time.perf_counter()
    sg, vg = xtal.native_setup(wl.xtal, e, wl.dmin, cut, device=0)
    t2 = time.perf_counter()
    lh = sh.lut.view(np.complex128)
    lg = sg.lut.view(np.complex128)
    umax = np.abs(lh).max()
    dev = np.abs(lg - lh) / umax
    bits = np.mean((lg.view(np.float64) == lh.view(np.float64)))
    _say(f"{name} {e:g} kV: {lh.size} cells, kernel {vg['lut_ms']:.3f} ms, bit-equal parts {bits:.4f}, "
         f"max |dU|/max|U| {dev.max():.2e}; whole setup on one host core {t1 - t0:.3f} s, "
         f"with the box on the GPU {t2 - t1:.3f} s")
    # list order, 000, wavelength: the host's
    assert np.array_equal(sg.hkl, sh.hkl)
    assert vg["mLambda"] == vh["mLambda"] and vg["DynUpz"] == vh["DynUpz"]
    # every cell bit-equal, and the list's Ucg and double-diffraction flags with it
    assert np.array_equal(lg.view(np.uint64), lh.view(np.uint64))
    assert np.array_equal(sg.ucg.view(np.uint64), sh.ucg.view(np.uint64))
    assert np.array_equal(sg.dbdiff, sh.dbdiff)


@pytest.mark.parametrize("name", FIXTURES)
def test_gpu_setup_against_reference(name):
    """The GPU-built setup against the reference's dump of the same crystal and energy: every
    quantity tests/test_setup_native.py compares, bit for bit."""
    from ctemsoft_amd import workloads, xtal
    fx = load_golden(name)
    wl = workloads.get(name.split("_")[0] + "_" + name.split("_")[1])
    st, v = xtal.native_setup(wl.xtal, float(fx["ekev"]) if "ekev" in fx.files else wl.ekev, wl.dmin,
                              (wl.weakcutoff, wl.cutoff, wl.sgcutoff), device=0)
    _say(f"{name}: {st.lut.size // 2} cells built on the GPU in {v['lut_ms']:.3f} ms")
    _compare(st, v, fx)


def test_gpu_setup_energy_series_against_reference():
    """The 21 energy bins of the Fe series (10-20 kV), each setup built on the GPU, against the
    reference's dumps."""
    from ctemsoft_amd import workloads, xtal
    wl = workloads.get("fe_c3w")
    arc = load_golden("fe_c3w_setups")
    ms = 0.0
    for i in range(int(arc["nenergy"])):
        view = {k[len(f"e{i}_"):]: arc[k] for k in arc.files if k.startswith(f"e{i}_")}
        st, v = xtal.native_setup(wl.xtal, float(view["ekev"]), wl.dmin, (wl.weakcutoff, wl.cutoff, wl.sgcutoff),
                                  device=0)
        ms += v["lut_ms"]
        _compare(st, v, view)
    _say(f"fe_c3w: {int(arc['nenergy'])} setups built on the GPU, LUT kernels {ms:.3f} ms in all")


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
