"""CPU-side checks of the HIP library: it loads, exports exactly the C-ABI that
include/*.h declare, reports a clean error without a GPU, and its host-side
functions (RoscaLambert k-vectors, Laue equivalents) equal the reference's."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

from conftest import REPO, lambert_dims, load_golden


def _declared(headers=("ctem_ebsd.h",)):
    txt = "".join(open(os.path.join(REPO, "include", h)).read() for h in headers)
    return sorted(set(re.findall(r"CTEM_API\s+[\w\s\*]+?\b(ctem_\w+)\s*\(", txt)))


def _all_headers():
    return tuple(sorted(f for f in os.listdir(os.path.join(REPO, "include")) if f.endswith(".h")))


def test_library_exports_every_declared_symbol():
    from ctemsoft_amd import detector, ebsd, ecp, mc, xtal
    L = ebsd.lib()
    decl = _declared()
    assert decl == sorted(ebsd.EXPORTS)
    assert _declared(("ctem_setup.h",)) == sorted(xtal.EXPORTS)
    assert _declared(("ctem_detector.h",)) == sorted(detector.EXPORTS)
    assert _declared(("ctem_ecp.h",)) == sorted(ecp.EXPORTS)
    assert _declared(("ctem_mc.h",)) == sorted(mc.EXPORTS)
    for name in _declared(_all_headers()):
        assert hasattr(L, name), name


def test_exports_match_nm():
    import subprocess
    from ctemsoft_amd import ebsd
    out = subprocess.run(["nm", "-D", "--defined-only", ebsd.LIBPATH], capture_output=True, text=True).stdout
    exported = sorted(l.split()[-1] for l in out.splitlines() if " T " in l and l.split()[-1].startswith("ctem_"))
    assert exported == _declared(_all_headers())


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


def test_laue_equiv_matches_reference_table():
    """ctem_laue_equiv for every Laue group 1..13 over the pixel box -16..16 squared against
    the reference's Apply2DLaueSymmetry table (Lambert.f90:1942-2046)."""
    from ctemsoft_amd import ebsd
    fx = load_golden("sym_tables")
    nb = int(fx["nbox"])
    bad = []
    for isym in range(1, 14):
        for iy in range(-nb, nb + 1):
            for ix in range(-nb, nb + 1):
                n = int(fx["laue_n"][isym - 1, iy + nb, ix + nb])
                ref = fx["laue_q"][isym - 1, iy + nb, ix + nb].reshape(12, 2)[:n].astype(np.int32)
                got = np.asarray(ebsd.laue_equiv(ix, iy, isym), np.int32).reshape(-1, 2)
                if got.shape != ref.shape or not np.array_equal(got, ref):
                    bad.append((isym, ix, iy))
    assert not bad, f"{len(bad)} pixels differ, first {bad[:5]}"


@pytest.mark.parametrize("hexgrid", [0, 1])
def test_kvectors_every_laue_group_match_reference(hexgrid):
    """ctem_kvectors for Laue groups 1..12 on the square and hexagonal grids against the
    reference's Calckvectors lists (kvectors.f90:421-542,839-947)."""
    from ctemsoft_amd import ebsd
    fx = load_golden("sym_tables")
    npx = int(fx["npx"])
    npy = int(round(2 * npx / np.sqrt(3.0))) if hexgrid else npx
    for isym in range(1, 13):
        ij, k, kn = ebsd.kvectors(npx, npy, isym, hexgrid, fx["dsm"], float(fx["mLambda"]))
        assert np.array_equal(ij, fx[f"kv{isym}_{hexgrid}_ij"]), isym
        assert np.array_equal(k, fx[f"kv{isym}_{hexgrid}_k"], equal_nan=True), isym
        assert np.array_equal(kn, fx[f"kv{isym}_{hexgrid}_kn"]), isym


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


def test_laue_group_of_every_space_group():
    """The host's Laue-group pick (isym, usehex) for space groups 1..230 against the reference's
    rule on its own tables (CTEMEBSDmaster.f90:249-282) and against the Laue classes of
    International Tables (space-group ranges of the 11 classes), with the fixtures' crystals."""
    from ctemsoft_amd import ebsd
    fx = load_golden("sym_tables")
    ranges = [(1, 2, 1), (3, 15, 2), (16, 74, 3), (75, 88, 4), (89, 142, 5), (143, 148, 6), (149, 167, 7),
              (168, 176, 8), (177, 194, 9), (195, 206, 10), (207, 230, 11)]
    for g in range(1, 231):
        isym, hexg = ebsd.laue_group(g)
        assert (isym, hexg) == (int(fx["laue_isym"][g - 1]), int(fx["laue_usehex"][g - 1])), g
        cls = next(c for a, b, c in ranges if a <= g <= b)
        assert isym == cls or (cls == 7 and isym == 12), g
    for name in ("ni_c0", "fe_c1", "al2o3_c2"):
        f = load_golden(name)
        assert ebsd.laue_group(int(f["sgnum"])) == (int(f["isym"]), int(f["usehex"]))
