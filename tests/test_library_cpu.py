"""CPU-side checks of the HIP library: it loads, exports exactly the C-ABI that
include/ctem_ebsd.h declares, reports a clean error without a GPU, and its host-side
functions (RoscaLambert k-vectors, Laue equivalents) equal the reference's."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

from conftest import REPO, lambert_dims, load_golden


def _declared():
    txt = open(os.path.join(REPO, "include", "ctem_ebsd.h")).read()
    return sorted(set(re.findall(r"CTEM_API\s+[\w\s\*]+?\b(ctem_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from ctemsoft_amd import ebsd
    L = ebsd.lib()
    decl = _declared()
    assert decl == sorted(ebsd.EXPORTS)
    for name in decl:
        assert hasattr(L, name), name


def test_exports_match_nm():
    import subprocess
    from ctemsoft_amd import ebsd
    out = subprocess.run(["nm", "-D", "--defined-only", ebsd.LIBPATH], capture_output=True, text=True).stdout
    exported = sorted(l.split()[-1] for l in out.splitlines() if " T " in l and l.split()[-1].startswith("ctem_"))
    assert exported == _declared()


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error path")
def test_no_gpu_error_is_clean():
    from ctemsoft_amd import ebsd
    with pytest.raises(ebsd.CtemError) as e:
        ebsd.Context(0)
    assert e.value.code == -2


@pytest.mark.parametrize("name", ["ni_c0", "fe_c1", "al2o3_c2"])
def test_kvectors_bit_exact(name):
    from ctemsoft_amd import ebsd
    fx = load_golden(name)
    npx, npy = lambert_dims(name, fx)
    ij, k, kn = ebsd.kvectors(npx, npy, int(fx["isym"]), int(fx["usehex"]), fx["dsm"], float(fx["mLambda"]))
    assert np.array_equal(ij, fx["kij"])
    assert np.array_equal(k, fx["kvec"])
    assert np.array_equal(kn, fx["kn"])


def test_kvectors_south_is_mirror():
    from ctemsoft_amd import ebsd
    fx = load_golden("fe_c1")
    ijn, kn, _ = ebsd.kvectors(30, 30, 11, 0, fx["dsm"], float(fx["mLambda"]), 1)
    ijs, ks, _ = ebsd.kvectors(30, 30, 11, 0, fx["dsm"], float(fx["mLambda"]), -1)
    assert np.array_equal(ijn, ijs)
    assert np.allclose(ks[:, :2], kn[:, :2], rtol=0, atol=1e-12) and np.allclose(ks[:, 2], -kn[:, 2], atol=1e-12)


@pytest.mark.parametrize("isym", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13])
def test_laue_equiv_matches_oracle(isym, oracle_mod):
    from ctemsoft_amd import ebsd
    for ipx, ipy in [(0, 0), (3, 1), (7, 4), (10, 10), (5, -3), (12, 0), (0, 9)]:
        assert np.array_equal(ebsd.laue_equiv(ipx, ipy, isym), oracle_mod.laue_equiv(ipx, ipy, isym))


def test_hex_kvectors_match_oracle(oracle_mod):
    from ctemsoft_amd import ebsd
    dsm = np.array([[0.4759, -0.23795, 0.0], [0.0, 0.41214, 0.0], [0.0, 0.0, 1.2991]])
    for isym in (6, 7, 8, 9, 12):
        a = ebsd.kvectors(20, 23, isym, 1, dsm, 0.0098)
        b = oracle_mod.kvectors(20, 23, isym, 1, dsm, 0.0098)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), isym


def test_pixel_owner_covers_the_master_array():
    """Every pixel of the square Lambert array is written by some direction of the m-3m wedge,
    and the owner is the last direction whose Laue equivalents include it."""
    from ctemsoft_amd import ebsd, fit
    fx = load_golden("ni_c0")
    ij, _, _ = ebsd.kvectors(12, 12, 11, 0, fx["dsm"], float(fx["mLambda"]))
    own = fit.pixel_owner(ij, 12, 12, 11)
    assert (own >= 0).all()
    i, j = ij[own[5 * 25 + 7]]
    q = ebsd.laue_equiv(int(i), int(j), 11)
    assert any((a + 12) + 25 * (b + 12) == 5 * 25 + 7 for a, b in q)


def test_target_setup_shares_factors_between_g_and_minus_g():
    from ctemsoft_amd import ebsd, fit
    fx = load_golden("ni_c0")
    st = ebsd.Setup.from_fixture(fx)
    tt = fit.perturbed_target_setup(st, 20264)
    d = st.lut_dim
    a = st.lut.view(np.complex128).reshape(d[2], d[1], d[0])
    b = tt.lut.view(np.complex128).reshape(d[2], d[1], d[0])
    m = np.abs(a) > 0
    f = np.where(m, b / np.where(m, a, 1), 0)
    assert np.allclose(f[m], f[::-1, ::-1, ::-1][m], rtol=1e-6)
    assert 0.005 < np.std(np.abs(f[m])) < 0.02
    assert np.allclose(tt.site_b, st.site_b * np.float32(1.05))
