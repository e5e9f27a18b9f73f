"""Pins for the oracle (oracle/ctem_oracle.c) before anything is compared against it.

1. Against the compiled reference (golden fixtures from tests/golden/make_golden.py):
   beam sets, DynMat, Lgh, k-vectors and the whole float32 master pattern.
2. Against mathematics that fixes the answer independently of the reference: the one-beam
   closed form, norm conservation without absorption, Hermitian PSD structure, a
   library matrix exponential and eigendecomposition, and brute-force RK4 integration.
"""
import numpy as np
import pytest
import scipy.linalg

from conftest import lambert_dims, load_golden

FIXTURES = ["ni_c0", "ni_c0w", "fe_c1", "fe_c1w", "al2o3_c2", "al2o3_c2w",
            "fe_c3w_10.0kV", "fe_c3w_15.0kV", "fe_c3w_20.0kV"]


def _split(fx, j, ns, nw):
    so = np.concatenate([[0], np.cumsum(ns)])
    wo = np.concatenate([[0], np.cumsum(nw)])
    return (fx["beams_strong"][so[j]:so[j + 1]], fx["beams_strongsg"][so[j]:so[j + 1]],
            fx["beams_weak"][wo[j]:wo[j + 1]], fx["beams_weaksg"][wo[j]:wo[j + 1]])


@pytest.mark.parametrize("name", FIXTURES)
def test_beam_sets_bit_exact(name, oracle_mod):
    fx = load_golden(name)
    cr = oracle_mod.Crystal(fx)
    ns, nw = fx["beams_nns"], fx["beams_nnw"]
    bad = []
    for j, k in enumerate(fx["beams_ik"]):
        st, ssg, wk, wsg = oracle_mod.select_beams(cr, fx["kvec"][k])
        rs, rssg, rw, rwsg = _split(fx, j, ns, nw)
        ok = (np.array_equal(st, rs) and np.array_equal(wk, rw)
              and np.array_equal(ssg.view(np.int32), rssg.view(np.int32))
              and np.array_equal(wsg.view(np.int32), rwsg.view(np.int32)))
        if not ok:
            bad.append(int(k))
    assert not bad, f"{len(bad)} directions with different beam sets: {bad[:10]}"


def test_southern_hemisphere_beam_sets_match_reference(oracle_mod):
    """The library's southern half of the corundum workload (al2o3_c2w; wave vectors from
    ctem_kvectors with hemisphere -1, the program itself computes the northern half only): the
    oracle's beam sets equal the reference's beam loop body on the same wave vectors
    (tests/golden/al2o3_c2w_south.npz), including the two directions with a weak beam ~3e-6
    from the Bragg condition. On those two the reference's CalcLgh3 gives a non-finite Lgh (its
    DynMat reaches ~1e4), on their northern twins a finite one."""
    fx = load_golden("al2o3_c2w")
    z = load_golden("al2o3_c2w_south")
    cr = oracle_mod.Crystal(fx)
    ns, nw = z["beams_nns"], z["beams_nnw"]
    for j in range(len(ns)):
        st, ssg, wk, wsg = oracle_mod.select_beams(cr, z["kvec"][j])
        rs, rssg, rw, rwsg = _split(z, j, ns, nw)
        assert np.array_equal(st, rs) and np.array_equal(wk, rw), int(z["beams_ik"][j])
        assert np.array_equal(ssg.view(np.int32), rssg.view(np.int32))
        assert np.array_equal(wsg.view(np.int32), rwsg.view(np.int32))
    assert not z["blow_south_lgh_finite"].any() and z["blow_north_lgh_finite"].all()
    assert (z["blow_south_min_sw"] < 1e-5).all() and (z["blow_south_dynmat_max"] > 1e4).all()


@pytest.mark.parametrize("name", FIXTURES)
def test_dynmat_and_lgh_match_reference(name, oracle_mod):
    fx = load_golden(name)
    cr = oracle_mod.Crystal(fx)
    for j, k in enumerate(fx["dyn_ik"]):
        st, ssg, wk, wsg = oracle_mod.select_beams(cr, fx["kvec"][k])
        dm = oracle_mod.dynmat(cr, st, ssg, wk, wsg)
        ref = fx[f"dyn{j}_dynmat"]
        assert np.array_equal(dm, ref), f"DynMat differs at direction {k}"
        L = oracle_mod.lgh(ref, fx["kn"][k], float(fx["dz"]), float(fx["thick"]), fx["lam"])
        Lr = fx[f"dyn{j}_lgh"]
        err = np.linalg.norm(L - Lr) / np.linalg.norm(Lr)
        # the reference's Lgh comes from zgeev/zgetri, whose error is ~cond(C) * eps
        assert err < 1e-12, (k, err)


@pytest.mark.parametrize("name", FIXTURES)
def test_kvectors_bit_exact(name, oracle_mod):
    fx = load_golden(name)
    npx, npy = lambert_dims(name, fx)
    ij, k, kn = oracle_mod.kvectors(npx, npy, int(fx["isym"]), int(fx["usehex"]), fx["dsm"],
                                    float(fx["mLambda"]))
    assert np.array_equal(ij, fx["kij"])
    assert np.array_equal(k, fx["kvec"])
    assert np.array_equal(kn, fx["kn"])


@pytest.mark.parametrize("name,nsample", [("ni_c0", None), ("ni_c0w", 12)])
def test_master_pattern_bit_exact(name, nsample, oracle_mod):
    """The float32 master pattern of the reference program, pixel for pixel."""
    fx = load_golden(name)
    cr = oracle_mod.Crystal(fx)
    sr = fx["master_sr"][0, 0]
    npx = (sr.shape[1] - 1) // 2
    nat = sum(len(s["orbit"]) for s in cr.sites)
    ks = range(len(fx["kij"])) if nsample is None else \
        np.random.default_rng(7).choice(len(fx["kij"]), nsample, replace=False)
    mism = 0
    for kk in ks:
        dots, _, _ = oracle_mod.master_direction(cr, fx["kvec"][kk], fx["kn"][kk], fx["lam"],
                                                  float(fx["dz"]), float(fx["thick"]))
        v = oracle_mod.svals_f32(dots[0], nat)
        for a, b in oracle_mod.laue_equiv(int(fx["kij"][kk, 0]), int(fx["kij"][kk, 1]),
                                          int(fx["isym"])):
            mism += int(np.float32(v).view(np.int32) != sr[b + npx, a + npx].view(np.int32))
    assert mism == 0


# ------------------------------- mathematical pins ------------------------------------
def _random_dynmat(n, rng, absorb=True, herm=False):
    a = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * 2.0
    if herm:
        a = 0.5 * (a + a.conj().T)
    d = np.diag(rng.uniform(-40, 40, n)).astype(complex)
    m = a + d
    if absorb:
        m = m + 1j * np.diag(np.full(n, 1.5))
    return m


def test_one_beam_closed_form(oracle_mod):
    fx = load_golden("ni_c0")
    cr = oracle_mod.Crystal(fx)
    k = int(np.flatnonzero(fx["beams_nns"] == 1)[0])
    k = int(fx["beams_ik"][k])
    st, ssg, wk, wsg = oracle_mod.select_beams(cr, fx["kvec"][k])
    assert len(st) == 1
    dm = oracle_mod.dynmat(cr, st, ssg, wk, wsg)
    lam, dz, thick, kn = fx["lam"].astype(float), float(fx["dz"]), float(fx["thick"]), fx["kn"][k]
    d = float(np.float32(2.0 * kn))
    z = dz * np.arange(1, len(lam) + 1)
    closed = dz / thick * np.sum(lam * np.exp(-4 * np.pi * z * dm[0, 0].imag / d))
    L = oracle_mod.lgh(dm, kn, dz, thick, lam)
    assert abs(L[0, 0] - closed) / closed < 1e-14
    assert abs(L[0, 0].imag) <= 1e-18 * abs(L[0, 0])


def test_lgh_against_scipy_expm_and_eig(oracle_mod):
    rng = np.random.default_rng(3)
    for n in (2, 5, 17, 40):
        m = _random_dynmat(n, rng)
        kn, dz, thick = 116.5, 0.5, 25.0
        lam = rng.uniform(0, 1, 50).astype(np.float32)
        L = oracle_mod.lgh(m, kn, dz, thick, lam)
        d = float(np.float32(2 * kn))
        E = scipy.linalg.expm(2j * np.pi * dz * m / d)
        psi = np.zeros(n, complex); psi[0] = 1
        ref = np.zeros((n, n), complex)
        for w in lam.astype(float):
            psi = E @ psi
            ref += w * np.outer(psi, psi.conj())
        ref *= dz / thick
        assert np.linalg.norm(L - ref) / np.linalg.norm(ref) < 1e-12
        # eigen route as in CalcLgh3: psi_k = C exp(2 pi i z W) C^-1 e1
        W, Cm = np.linalg.eig(m / d)
        alpha = np.linalg.solve(Cm, np.eye(n)[:, 0])
        ref2 = np.zeros((n, n), complex)
        for i, w in enumerate(lam.astype(float)):
            p = Cm @ (np.exp(2j * np.pi * dz * (i + 1) * W) * alpha)
            ref2 += w * np.outer(p, p.conj())
        ref2 *= dz / thick
        assert np.linalg.norm(L - ref2) / np.linalg.norm(ref2) < 1e-10


def test_lgh_hermitian_psd_and_trace(oracle_mod):
    rng = np.random.default_rng(5)
    n = 12
    m = _random_dynmat(n, rng, absorb=False, herm=True)
    lam = rng.uniform(0, 1, 30).astype(np.float32)
    L = oracle_mod.lgh(m, 116.0, 1.0, 30.0, lam)
    assert np.allclose(L, L.conj().T, atol=1e-15)
    assert np.linalg.eigvalsh(L).min() > -1e-14
    # no absorption, Hermitian DynMat: |psi_k| = 1, so tr L = (dz/thick) sum(lam)
    assert abs(np.trace(L).real - np.sum(lam.astype(float)) / 30.0) < 1e-13


def test_lgh_against_rk4(oracle_mod):
    """Brute force: integrate d psi/dz = 2 pi i M' psi with many RK4 steps."""
    rng = np.random.default_rng(11)
    n = 3
    m = _random_dynmat(n, rng)
    kn, dz, thick = 100.0, 1.0, 4.0
    lam = np.array([0.4, 0.3, 0.2, 0.1], np.float32)
    d = float(np.float32(2 * kn))
    A = 2j * np.pi * m / d
    psi = np.zeros(n, complex); psi[0] = 1
    h = dz / 4000
    ref = np.zeros((n, n), complex)
    for w in lam.astype(float):
        for _ in range(4000):
            k1 = A @ psi; k2 = A @ (psi + 0.5 * h * k1); k3 = A @ (psi + 0.5 * h * k2); k4 = A @ (psi + h * k3)
            psi = psi + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        ref += w * np.outer(psi, psi.conj())
    ref *= dz / thick
    L = oracle_mod.lgh(m, kn, dz, thick, lam)
    assert np.linalg.norm(L - ref) / np.linalg.norm(ref) < 1e-10


@pytest.mark.parametrize("isym,order", [(2, 2), (3, 4), (4, 4), (5, 8), (11, 8)])
def test_laue_equivalents_form_orbits(oracle_mod, isym, order):
    """Square-grid Laue operations: the image set of a generic pixel has the group order
    and is the same set from every member (closure)."""
    q = oracle_mod.laue_equiv(7, 3, isym)
    assert len(q) == order and len({tuple(x) for x in q}) == order
    s = {tuple(x) for x in q}
    for a, b in q:
        assert {tuple(x) for x in oracle_mod.laue_equiv(int(a), int(b), isym)} == s


def test_calcsg_ewald_identity(oracle_mod):
    """With the foil normal along k the excitation error is (k^2 - |k+g|^2)/(2|k+g|cos a),
    cos a = (k+g).k/(|k+g||k|); check against a float64 evaluation."""
    rmt = np.diag([1 / 0.3**2] * 3)
    k = np.array([0.3, 1.1, 38.0])
    for g in ([1, 1, 0], [2, 0, 0], [1, -1, 2], [0, 0, -2]):
        s = oracle_mod.calcsg(g, k, k, rmt)
        kpg = (k + g) / 0.3; kk = k / 0.3
        cosa = kpg @ kk / (np.linalg.norm(kpg) * np.linalg.norm(kk))
        ref = (kk @ kk - kpg @ kpg) / (2 * np.linalg.norm(kpg) * cosa)
        assert abs(s - ref) <= 2e-6 * max(1.0, abs(ref))


def test_corundum_beam_sets_reach_beyond_lut_box(oracle_mod):
    """The corundum list holds family members with |h+k| up to twice the LUT box, so beam pairs
    of one direction have g-h outside the box; the bit-exact DynMat comparison above then
    exercises the reference's column-major (aliased) LUT reads."""
    fx = load_golden("al2o3_c2w")
    cr = oracle_mod.Crystal(fx)
    h = fx["hkl"].astype(np.int64)
    lo = fx["lut_lo"].astype(np.int64)
    hi = lo + np.array(fx["lut"].shape[::-1]) - 1
    st, ssg, wk, wsg = oracle_mod.select_beams(cr, fx["kvec"][fx["dyn_ik"][0]])
    g = h[np.concatenate([st, wk]) - 1]
    d = g[:len(st), None, :] - g[None, :, :]
    outside = ((d < lo) | (d > hi)).any(-1)
    assert outside.sum() > 100


def _sym_cases():
    fx = load_golden("sym_tables")
    nb, npx = int(fx["nbox"]), int(fx["npx"])
    return fx, nb, npx, int(round(2 * npx / np.sqrt(3.0)))


def test_laue_equivalents_match_reference_table(oracle_mod):
    """Apply2DLaueSymmetry (Lambert.f90:1942-2046) for every Laue group 1..13 over the pixel
    box -16..16 squared, against the reference's own table (tests/golden/sym_tables.npz)."""
    fx, nb, _, _ = _sym_cases()
    bad = []
    for isym in range(1, 14):
        for iy in range(-nb, nb + 1):
            for ix in range(-nb, nb + 1):
                n = int(fx["laue_n"][isym - 1, iy + nb, ix + nb])
                ref = fx["laue_q"][isym - 1, iy + nb, ix + nb].reshape(12, 2)[:n].astype(np.int32)
                got = np.asarray(oracle_mod.laue_equiv(ix, iy, isym), np.int32).reshape(-1, 2)
                if got.shape != ref.shape or not np.array_equal(got, ref):
                    bad.append((isym, ix, iy))
    assert not bad, f"{len(bad)} pixels differ, first {bad[:5]}"


@pytest.mark.parametrize("hexgrid", [0, 1])
@pytest.mark.parametrize("isym", list(range(1, 13)))
def test_kvectors_every_laue_group_match_reference(isym, hexgrid, oracle_mod):
    """Calckvectors/AddkVector (kvectors.f90:421-542,839-947) for every Laue group on the square
    and hexagonal grids against the reference's lists (hexagonal grids take npy = npyhex, as
    the program passes it; points outside the hexagon are NaN in both)."""
    fx, _, npx, npyhex = _sym_cases()
    npy = npyhex if hexgrid else npx
    ij, k, kn = oracle_mod.kvectors(npx, npy, isym, hexgrid, fx["dsm"], float(fx["mLambda"]))
    assert np.array_equal(ij, fx[f"kv{isym}_{hexgrid}_ij"])
    assert np.array_equal(k, fx[f"kv{isym}_{hexgrid}_k"], equal_nan=True)
    assert np.array_equal(kn, fx[f"kv{isym}_{hexgrid}_kn"])


@pytest.mark.parametrize("name", ["fe_c1w", "ni_c0w", "al2o3_c2w", "fe_c3w_10.0kV"])
def test_mixed_form_truncation_at_its_scaling_threshold(name, oracle_mod):
    """The mixed form scales A until ||A / 2^s||_1 <= 5/4 and takes the degree-8 Taylor polynomial
    of exp - I there (F' = F^2 + 2F per squaring). In FP64 arithmetic, without the split
    products' rounding, that truncation moves the intensities Re sum Lgh S by less than 1e-5
    relative to the exact exponential (scipy) on sampled directions of each crystal, while one
    squaring fewer (threshold 3) would move them by up to ~1e-3."""
    import math
    fx = load_golden(name)
    cr = oracle_mod.Crystal(fx)
    dz, thick = float(fx["dz"]), float(fx["thick"])
    lam = np.asarray(fx["lam"], float)
    rng = np.random.default_rng(2)

    def taylor8(B):
        F = np.zeros_like(B)
        P = np.eye(len(B), dtype=complex)
        for k in range(1, 9):
            P = P @ B / k
            F = F + P
        return F

    def dots(E, S):
        psi = np.zeros(len(E), complex)
        psi[0] = 1
        L = np.zeros(E.shape, complex)
        for lk in lam:
            psi = E @ psi
            L += lk * np.outer(psi, psi.conj())
        return float(np.real(np.sum(L * S))) * dz / thick

    worst = {1.25: 0.0, 3.0: 0.0}
    for k in rng.choice(len(fx["kvec"]), 4, replace=False):
        st, ssg, wk, wsg = oracle_mod.select_beams(cr, fx["kvec"][k])
        A = 2j * np.pi * dz * oracle_mod.dynmat(cr, st, ssg, wk, wsg) / np.float32(2 * float(fx["kn"][k]))
        S = oracle_mod.sgh(cr, st, 0)
        ref = dots(scipy.linalg.expm(A), S)
        nrm = np.abs(A).sum(axis=0).max()
        for theta in worst:
            s = max(0, math.ceil(math.log2(nrm / theta)))
            F = taylor8(A / 2.0 ** s)
            for _ in range(s):
                F = F @ F + 2 * F
            worst[theta] = max(worst[theta], abs(dots(np.eye(len(A)) + F, S) - ref) / abs(ref))
    print(f"{name}: truncation at threshold 5/4 {worst[1.25]:.2e}, at 3 {worst[3.0]:.2e}")
    assert worst[1.25] < 5e-6
    assert worst[3.0] > 10 * worst[1.25]


@pytest.mark.parametrize("name,count", [("fe_c1w", 12), ("fe_c3w_10.0kV", 6)])
def test_oracle_dots_match_reference_on_many_directions(name, count, oracle_mod):
    """The oracle's intensities (extended-precision Lgh) against the reference's own CalcLgh3 on
    seeded directions of tests/golden/ref_dots.npz (the reference's Lgh contracted with the
    per-site S_p), within 1e-10 relative."""
    from concurrent.futures import ThreadPoolExecutor
    fx = load_golden(name)
    z = load_golden("ref_dots")
    cr = oracle_mod.Crystal(fx)
    ks, ref = z[f"{name}_ik"][:count], z[f"{name}_dots"][:count]

    def one(k):
        return oracle_mod.master_direction(cr, fx["kvec"][k], fx["kn"][k], fx["lam"], float(fx["dz"]),
                                           float(fx["thick"]))[0]
    with ThreadPoolExecutor(8) as ex:
        got = np.array(list(ex.map(one, ks)))
    rel = np.abs(got - ref) / np.abs(ref)
    print(f"{name}: oracle vs reference dots on {len(ks)} directions, max rel {rel.max():.2e}")
    assert rel.max() <= 1e-10
