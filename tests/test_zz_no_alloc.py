"""ctem_calclgh3_static -- the allocation-free variant of the CalcLgh3 drop-in, whose work space
is a fixed device array per GPU -- called through the C-ABI against ctem_calclgh3 on the same
inputs, two calls in a row with the second one smaller, and against the reference's Lgh."""
import ctypes as C

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _call(fn, ctx, dm, thick, kn, dz, lam):
    from ctemsoft_amd import ebsd
    d = np.asfortranarray(dm, np.complex128)
    n = d.shape[0]
    out = np.zeros((n, n), np.complex128, order="F")
    lam = np.ascontiguousarray(lam, np.float32)
    rc = getattr(ebsd.lib(), fn)(ctx._h, d.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
                                 float(thick), float(kn), n, 1, float(dz), lam.ctypes.data_as(C.c_void_p), len(lam))
    return rc, np.array(out)


def test_calclgh3_static_matches_calclgh3():
    from ctemsoft_amd import ebsd
    ctx = ebsd.Context(0)
    cases = []
    fx = load_golden("fe_c1w")  # n ~ 210, 50 depths
    k = int(fx["dyn_ik"][0])
    cases.append(("fe_c1w", fx["dyn0_dynmat"], fx["dyn0_lgh"], float(fx["thick"]), float(fx["kn"][k]),
                  float(fx["dz"]), fx["lam"]))
    fx = load_golden("ni_c0w")  # smaller: n ~ 90, 20 depths
    k = int(fx["dyn_ik"][1])
    cases.append(("ni_c0w", fx["dyn1_dynmat"], fx["dyn1_lgh"], float(fx["thick"]), float(fx["kn"][k]),
                  float(fx["dz"]), fx["lam"]))
    for name, dm, Lr, thick, kn, dz, lam in cases:
        print(f"{name}: n = {dm.shape[0]}, {len(lam)} depths", flush=True)
        rc, Ls = _call("ctem_calclgh3_static", ctx, dm, thick, kn, dz, lam)
        assert rc == 0, ebsd.lib().ctem_last_error()
        rc, La = _call("ctem_calclgh3", ctx, dm, thick, kn, dz, lam)
        assert rc == 0, ebsd.lib().ctem_last_error()
        diff = np.abs(Ls - La).max() / np.abs(La).max()
        err = np.linalg.norm(Ls - Lr) / np.linalg.norm(Lr)
        print(f"  static vs allocating: max rel {diff:.1e}; static vs reference Lgh {err:.1e}", flush=True)
        assert diff <= 1e-14
        assert err < 1e-10
    # beyond the fixed work space: a clear error, no work done
    big = np.eye(300, dtype=np.complex128)
    rc, _ = _call("ctem_calclgh3_static", ctx, big, 25.0, 100.0, 0.5, np.ones(50, np.float32) / 50)
    assert rc == -4 and b"at most 256" in ebsd.lib().ctem_last_error()
    ctx.close()
