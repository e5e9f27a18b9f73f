"""ctypes wrapper of the oracle (oracle/_ref/liboracle.so, built from ctem_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by bench.py's
CPU-baseline leg, always as the checker and never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_ref", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        L = C.CDLL(path)
        P = C.c_void_p
        L.orc_calcsg.restype = C.c_float
        L.orc_calcsg.argtypes = [P, P, P, P]
        L.orc_select_beams.argtypes = [C.c_int, P, P, P, C.c_double, C.c_float, C.c_float,
                                       C.c_float, P, P, P, P, P, P, P, P]
        L.orc_dynmat.argtypes = [C.c_int, P, P, C.c_int, P, P, P, P, P, P, C.c_double,
                                 C.c_float, P]
        L.orc_sgh.argtypes = [C.c_int, P, P, P, C.c_int32, C.c_float, C.c_int, P, P]
        L.orc_lgh.argtypes = [C.c_int, P, C.c_double, C.c_double, C.c_double, C.c_int, P, P]
        L.orc_lgh_dot_sgh.restype = C.c_double
        L.orc_lgh_dot_sgh.argtypes = [C.c_int, P, P]
        L.orc_kvectors.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, P, C.c_double, C.c_int,
                                   P, P, P]
        L.orc_laue_equiv.argtypes = [C.c_int, C.c_int, C.c_int, P]
        F, I = C.c_float, C.c_int
        L.orc_ebsd_detector.argtypes = [F, F, F, I, I, F, F, F, I, I, I, P, I, I, P, P, P]
        L.orc_ebsd_pattern.argtypes = [I, I, P, P, I, I, I, I, I, P, P, C.c_double, P]
        L.orc_poisson_noise.argtypes = [P, I, I, P]
        L.orc_eu2qu.argtypes = [P, P]
        L.orc_apply_axisangle.argtypes = [I, P, P]
        L.orc_site_sum.argtypes = [C.c_int64, I, P, P]
        L.orc_bin_pattern.argtypes = [I, I, I, I, F, P, P]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Crystal:
    """Per-energy setup arrays, as dumped from the reference (or built by the host)."""

    def __init__(self, fx, lut_from=None):
        """lut_from: a fixture of the same crystal and energy holding `lut_full`, the LUT over
        the box of every difference of two list entries (hexagonal axes), used in place of the
        reference-sized LUT."""
        self.hkl = np.ascontiguousarray(fx["hkl"], np.int32)
        ucg = np.asarray(fx["ucg"], np.complex128)
        self.ucg = np.ascontiguousarray(np.stack([ucg.real, ucg.imag], -1).ravel())
        lut = np.asarray(fx["lut"], np.complex128)  # [l][k][h] = Fortran LUT(h,k,l)
        lo = np.asarray(fx["lut_lo"], np.int32).copy()
        dbd = np.asarray(fx["dbdiff"])
        h, k, l = (self.hkl - lo).T
        self.dbdiff_ref = np.ascontiguousarray(dbd[l, k, h].astype(np.uint8))
        if lut_from is not None:
            assert np.array_equal(np.asarray(lut_from["hkl"]), np.asarray(fx["hkl"]))
            lut = np.asarray(lut_from["lut_full"], np.complex128)
            lo = np.asarray(lut_from["lut_full_lo"], np.int32).copy()
        self.lut_dim = np.array(lut.shape[::-1], np.int32)
        self.lut_lo = lo
        self.lut = np.ascontiguousarray(np.stack([lut.real, lut.imag], -1).ravel())
        self.mlambda = float(fx["mLambda"])
        self.dynupz = float(fx["DynUpz"])
        self.weakcutoff, self.cutoff, self.sgcutoff = (float(v) for v in fx["cutoffs"])
        self.rmt = np.ascontiguousarray(fx["rmt"], np.float64).ravel()
        self.dsm = np.ascontiguousarray(fx["dsm"], np.float64).ravel()
        self.nsite = int(fx["nsite"])
        self.sites = [dict(Z=int(fx[f"site{i}_Z"]), pos=np.asarray(fx[f"site{i}_pos"]),
                           orbit=np.ascontiguousarray(fx[f"site{i}_orbit"], np.float64))
                      for i in range(self.nsite)]
        self.nref = len(self.hkl)


def calcsg(g, k, fn, rmt):
    g = np.asarray(g, np.float32); k = np.asarray(k, np.float32); fn = np.asarray(fn, np.float32)
    return lib().orc_calcsg(_p(g), _p(k), _p(fn), _p(np.asarray(rmt, np.float64)))


def select_beams(cr: Crystal, kvec):
    n = cr.nref
    st = np.zeros(n, np.int32); ssg = np.zeros(n, np.float32)
    wk = np.zeros(n, np.int32); wsg = np.zeros(n, np.float32)
    ns = C.c_int32(); nw = C.c_int32()
    kv = np.ascontiguousarray(kvec, np.float64)
    lib().orc_select_beams(n, _p(cr.hkl), _p(cr.ucg), _p(cr.dbdiff_ref), cr.mlambda,
                           cr.weakcutoff, cr.cutoff, cr.sgcutoff, _p(cr.rmt), _p(kv), _p(st),
                           _p(ssg), C.byref(ns), _p(wk), _p(wsg), C.byref(nw))
    return st[:ns.value].copy(), ssg[:ns.value].copy(), wk[:nw.value].copy(), wsg[:nw.value].copy()


def dynmat(cr: Crystal, strong, strongsg, weak, weaksg):
    n = len(strong)
    out = np.zeros(2 * n * n)
    s = np.ascontiguousarray(strong, np.int32); ssg = np.ascontiguousarray(strongsg, np.float32)
    w = np.ascontiguousarray(weak, np.int32); wsg = np.ascontiguousarray(weaksg, np.float32)
    lib().orc_dynmat(n, _p(s), _p(ssg), len(w), _p(w), _p(wsg), _p(cr.hkl), _p(cr.lut),
                     _p(cr.lut_lo), _p(cr.lut_dim), cr.mlambda, cr.dynupz, _p(out))
    return out.view(np.complex128).reshape(n, n)


def sgh(cr: Crystal, strong, site: int):
    n = len(strong)
    out = np.zeros(2 * n * n)
    s = np.ascontiguousarray(strong, np.int32)
    st = cr.sites[site]
    lib().orc_sgh(n, _p(s), _p(cr.hkl), _p(cr.rmt), st["Z"], float(st["pos"][4]),
                  len(st["orbit"]), _p(st["orbit"]), _p(out))
    return out.view(np.complex128).reshape(n, n)


def lgh(dm, kn, dz, thick, lam):
    n = dm.shape[0]
    d = np.ascontiguousarray(dm, np.complex128)
    lam = np.ascontiguousarray(lam, np.float32)
    out = np.zeros(2 * n * n)
    lib().orc_lgh(n, _p(d.view(np.float64)), kn, dz, thick, len(lam), _p(lam), _p(out))
    return out.view(np.complex128).reshape(n, n)


def lgh_dot_sgh(L, S):
    n = L.shape[0]
    a = np.ascontiguousarray(L, np.complex128); b = np.ascontiguousarray(S, np.complex128)
    return lib().orc_lgh_dot_sgh(n, _p(a.view(np.float64)), _p(b.view(np.float64)))


def kvectors(npx, npy, isym, usehex, dsm, mlambda):
    cap = (2 * npx + 1) * (2 * max(npx, npy) + 1) * 2 + 8
    ij = np.zeros(2 * cap, np.int32); k = np.zeros(3 * cap); kn = np.zeros(cap)
    n = lib().orc_kvectors(npx, npy, isym, int(usehex), _p(np.ascontiguousarray(dsm, np.float64).ravel()),
                           mlambda, cap, _p(ij), _p(k), _p(kn))
    if n < 0:
        raise RuntimeError(f"orc_kvectors failed ({n})")
    return ij[:2 * n].reshape(n, 2).copy(), k[:3 * n].reshape(n, 3).copy(), kn[:n].copy()


def laue_equiv(ipx, ipy, isym):
    q = np.zeros(24, np.int32)
    ne = lib().orc_laue_equiv(ipx, ipy, isym, _p(q))
    return q[:2 * ne].reshape(ne, 2).copy()


def svals_f32(dot: float, nat_total: int) -> np.float32:
    """svals(ip) = real(sum(Lgh*Sgh)) stored in float32, then divided by float(sum(nat))
    in float32 (CTEMEBSDmaster.f90:456-459)."""
    return np.float32(np.float32(dot) / np.float32(nat_total))


def master_direction(cr: Crystal, kvec, kn, lam, dz, thick):
    """Oracle for one incident direction: svals per site in double (before the float32
    store) and the beam set."""
    st, ssg, wk, wsg = select_beams(cr, kvec)
    dm = dynmat(cr, st, ssg, wk, wsg)
    L = lgh(dm, kn, dz, thick, lam)
    dots = np.array([lgh_dot_sgh(L, sgh(cr, st, p)) for p in range(cr.nsite)])
    return dots, st, wk


# ---- EBSD detector patterns (CTEMEBSD.f90) ----------------------------------------------------

def ebsd_detector(geom: dict, accum_e, ehistmin: float, ebinsize: float, mcsig: float):
    """The program's detector set-up: (rg [3][ny][nx], wdet [ny][nx][nE], Emin, Emax, num_el,
    prefactor). geom: &EBSDdata entries; accum_e [2nsy+1][2nsx+1][nE] int32."""
    acc = np.ascontiguousarray(accum_e, np.int32)
    ny2, nx2, ne = acc.shape
    f32 = np.float32
    emin = int(np.floor(abs((float(f32(geom["energymin"])) - ehistmin) / ebinsize) + 0.5)
               * np.sign((float(f32(geom["energymin"])) - ehistmin) / ebinsize)) + 1
    emax = int(np.floor(abs((float(f32(geom["energymax"])) - ehistmin) / ebinsize) + 0.5)
               * np.sign((float(f32(geom["energymax"])) - ehistmin) / ebinsize)) + 1
    emin, emax = min(max(emin, 1), ne), min(max(emax, 1), ne)
    nx, ny = int(geom["numsx"]), int(geom["numsy"])
    rg = np.zeros(3 * nx * ny, np.float32)
    wdet = np.zeros(nx * ny * ne, np.float32)
    nel = C.c_int64()
    lib().orc_ebsd_detector(geom["L"], geom["thetac"], geom["delta"], nx, ny, geom["xpc"], geom["ypc"],
                            float(f32(mcsig)), ne, (nx2 - 1) // 2, (ny2 - 1) // 2, _p(acc), emin, emax, _p(rg),
                            _p(wdet), C.byref(nel))
    pref = 0.25 * 6.241e18 * float(geom["beamcurrent"]) * float(geom["dwelltime"]) * 1.0e-15 / float(nel.value)
    return rg, wdet, emin, emax, int(nel.value), pref


def ebsd_pattern(geom: dict, det, sr, quat) -> np.ndarray:
    """One noise-free pattern [ny][nx] from the site-summed master sr [nE][2npy+1][2npx+1]."""
    rg, wdet, emin, emax, _, pref = det
    s = np.ascontiguousarray(sr, np.float32)
    ne, ny2, nx2 = s.shape
    nx, ny = int(geom["numsx"]), int(geom["numsy"])
    q = np.ascontiguousarray(quat, np.float32)
    out = np.zeros(nx * ny, np.float32)
    lib().orc_ebsd_pattern(nx, ny, _p(rg), _p(wdet), ne, emin, emax, (nx2 - 1) // 2, (ny2 - 1) // 2, _p(s), _p(q),
                           pref, _p(out))
    return out.reshape(ny, nx)


def poisson_noise(pattern, state) -> np.ndarray:
    p = np.array(pattern, np.float32)
    ny, nx = p.shape
    lib().orc_poisson_noise(_p(p), nx, ny, _p(state))
    return p


def eu2qu_deg(eul) -> np.ndarray:
    e = (np.asarray(eul, np.float32).reshape(-1, 3) * np.float32(0.0174533)).astype(np.float32)
    q = np.zeros((len(e), 4), np.float32)
    for i in range(len(e)):
        ei = np.ascontiguousarray(e[i])
        qi = np.zeros(4, np.float32)
        lib().orc_eu2qu(_p(ei), _p(qi))
        q[i] = qi
    return q


def apply_axisangle(quats, ax) -> np.ndarray:
    q = np.array(quats, np.float32).reshape(-1, 4)
    a = np.ascontiguousarray(ax, np.float32)
    lib().orc_apply_axisangle(len(q), _p(a), _p(q))
    return q


def site_sum(srtmp) -> np.ndarray:
    s = np.ascontiguousarray(srtmp, np.float32)
    out = np.zeros(s.shape[1:], np.float32)
    lib().orc_site_sum(out.size, s.shape[0], _p(s), _p(out))
    return out


def bin_pattern(pattern, binning: int, gamma: bool, gammavalue: float) -> np.ndarray:
    p = np.ascontiguousarray(pattern, np.float32)
    ny, nx = p.shape
    out = np.zeros((ny // binning, nx // binning), np.float32)
    lib().orc_bin_pattern(nx, ny, binning, int(gamma), float(gammavalue), _p(p), _p(out))
    return out


# ----------------------------------------------------------------------------------------
# electron channeling patterns (src_programs/CTEMECP.f90)

class EcpCrystal:
    """The set-up the channeling-pattern beam loop reads (reflection list, LUT, sites), from the
    arrays of the reference's set-up dump (oracle/ref_ecp.f90) or of the native set-up."""

    def __init__(self, hkl, ucg, dbdiff_ref, lut, lut_lo, mlambda, dynupz, cutoffs, rmt, sites):
        self.hkl = np.ascontiguousarray(hkl, np.int32).reshape(-1, 3)
        u = np.asarray(ucg, np.complex128)
        self.ucg = np.ascontiguousarray(np.stack([u.real, u.imag], -1).ravel())
        self.dbdiff_ref = np.ascontiguousarray(dbdiff_ref, np.uint8)
        lutc = np.asarray(lut, np.complex128)  # [l][k][h]
        self.lut_dim = np.array(lutc.shape[::-1], np.int32)
        self.lut_lo = np.ascontiguousarray(lut_lo, np.int32)
        self.lut = np.ascontiguousarray(np.stack([lutc.real, lutc.imag], -1).ravel())
        self.mlambda, self.dynupz = float(mlambda), float(dynupz)
        self.weakcutoff, self.cutoff, self.sgcutoff = (float(np.float32(c)) for c in cutoffs)
        self.rmt = np.ascontiguousarray(rmt, np.float64).ravel()
        self.sites = sites  # dict(Z, B, occ, orbit [n][3])
        self.nref = len(self.hkl)
        self.nat = int(sum(len(s["orbit"]) for s in sites))


def select_beams_fn(cr, kvec, fn):
    """orc_select_beams_fn: the beam sets with the foil normal fn (None: FN = k)."""
    n = cr.nref
    st = np.zeros(n, np.int32); ssg = np.zeros(n, np.float32)
    wk = np.zeros(n, np.int32); wsg = np.zeros(n, np.float32)
    ns = C.c_int32(); nw = C.c_int32()
    kv = np.ascontiguousarray(kvec, np.float64)
    f = None if fn is None else _p(np.ascontiguousarray(fn, np.float32))
    L = lib()
    L.orc_select_beams_fn.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_float,
                                      C.c_float, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_select_beams_fn(n, _p(cr.hkl), _p(cr.ucg), _p(cr.dbdiff_ref), cr.mlambda, cr.weakcutoff, cr.cutoff,
                          cr.sgcutoff, _p(cr.rmt), _p(kv), f, _p(st), _p(ssg), C.byref(ns), _p(wk), _p(wsg),
                          C.byref(nw))
    return st[:ns.value].copy(), ssg[:ns.value].copy(), wk[:nw.value].copy(), wsg[:nw.value].copy()


def sgh_ecp(cr, strong):
    """orc_sgh_ecp: the channeling structure matrix (all sites, occupancies) of the strong beams."""
    n = len(strong)
    out = np.zeros(2 * n * n)
    s = np.ascontiguousarray(strong, np.int32)
    z = np.array([st["Z"] for st in cr.sites], np.int32)
    b = np.array([st["B"] for st in cr.sites], np.float32)
    occ = np.array([st["occ"] for st in cr.sites], np.float32)
    norb = np.array([len(st["orbit"]) for st in cr.sites], np.int32)
    orb = np.ascontiguousarray(np.concatenate([np.asarray(st["orbit"], np.float64).ravel() for st in cr.sites]))
    L = lib()
    L.orc_sgh_ecp.argtypes = [C.c_int] + [C.c_void_p] * 3 + [C.c_int] + [C.c_void_p] * 6
    L.orc_sgh_ecp(n, _p(s), _p(cr.hkl), _p(cr.rmt), len(cr.sites), _p(z), _p(b), _p(occ), _p(norb), _p(orb),
                  _p(out))
    return out.view(np.complex128).reshape(n, n)


def matrix_exponential_pade(A, z0):
    """MatrixExponential with 'Pade' (math.f90:187-315): the diagonal Pade approximant of degree q
    with 2^j scaling, q and j from ilev = nint(log10(max |a_ij| z0)) + 3; the inverse of the
    denominator through LAPACK (numpy's), then j squarings."""
    q_tab, j_tab = (2, 3, 4, 4, 4, 4), (0, 0, 1, 5, 8, 11)
    modA = float(np.max(np.abs(A) * z0))
    if modA > 1000.0:
        raise ValueError("MatrixExponential: ||A|| > 1000")
    x = float(np.log10(modA)) if modA > 0 else -100.0
    ilev = int(np.copysign(np.floor(abs(x) + 0.5), x)) + 3  # nint: halves away from zero
    ilev = min(max(ilev, 1), 6)
    q, j = q_tab[ilev - 1], j_tab[ilev - 1]
    B = (A * z0) / 2.0 ** j
    C_ = B.copy()
    n = len(A)
    N = np.eye(n, dtype=np.complex128)
    D = N.copy()
    pref, sgn = 1.0, -1.0
    for icnt in range(1, q + 1):
        pref = pref * float(q - icnt + 1) / float(icnt) / float(2 * q - icnt + 1)
        N = N + pref * C_
        D = D + sgn * pref * C_
        sgn = -sgn
        C_ = C_ @ B
    E = np.linalg.inv(D) @ N
    for _ in range(j):
        E = E @ E
    return E


def ecp_lgh_scatmat(dm, mlambda, thick, zintstep, exact=False):
    """CalcLghSM (CTEMECP.f90:559-618): Azz = expm(i pi mLambda zintstep DynMat) (the reference's
    Pade rule, or with exact=True scipy's expm), psi_i = Azz^i e_1, Lgh(t) = sum_{i <= tvals(t)}
    psi_i psi_i^H with tvals = nint(thick/zintstep) met in order; [nt][n][n]."""
    import scipy.linalg
    thick = np.asarray(thick, np.float32)
    zs = np.float32(zintstep)
    Minp = dm * complex(0.0, np.pi * mlambda)
    Azz = scipy.linalg.expm(Minp * float(zs)) if exact else matrix_exponential_pade(Minp, float(zs))
    n = len(dm)
    nt = len(thick)
    tvals = [int(np.floor(abs(float(v)) + 0.5) * np.sign(float(v))) for v in thick / zs]
    numt = int(np.floor(float(np.max(thick) / zs) + 0.5))
    psi = np.zeros(n, np.complex128)
    psi[0] = 1.0
    lsum = np.zeros((n, n), np.complex128)
    out = np.zeros((nt, n, n), np.complex128)
    tval = 0
    for i in range(1, numt + 1):
        psi = Azz @ psi
        lsum = lsum + np.outer(psi, np.conj(psi))
        if tval < nt and i == tvals[tval]:
            out[tval] = lsum
            tval += 1
    return out


def ecp_lgh_blochwv(dm, kn, thick, round_q=True):
    """CalcLgh (CTEMECP.f90:472-534): the Bloch waves of DynMat (numpy's zgeev), W / fl32(2 kn),
    Ijk(i,j) = conj(a_i) (1 - exp(-q)) / q a_j with a = CGinv(:, 1) and
    q = 2 pi thick cmplx(Im W_i + Im W_j, Re W_i - Re W_j) -- the cmplx of default kind rounds both
    parts to float32 (round_q=False keeps them in double) -- and Lgh = CG Ijk^T CG^H; [nt][n][n]."""
    W, CG = np.linalg.eig(np.asarray(dm, np.complex128))
    CGinv = np.linalg.inv(CG)
    W = W / complex(float(np.float32(2.0 * kn)), 0.0)
    a = CGinv[:, 0]
    out = []
    for t in np.asarray(thick, np.float32):
        re = W.imag[:, None] + W.imag[None, :]
        im = W.real[:, None] - W.real[None, :]
        if round_q:
            re = re.astype(np.float32).astype(np.float64)
            im = im.astype(np.float32).astype(np.float64)
        q = (2.0 * np.pi * float(t)) * (re + 1j * im)
        Ijk = ((np.conj(a)[:, None] * (1.0 - np.exp(-q))) / q) * a[None, :]
        out.append(CG @ Ijk.T @ np.conj(CG).T)
    return np.array(out)


def ecp_direction(cr, kvec, kn, fn, mode, thick, zintstep=1.0, exact=False):
    """One incident direction of the channeling-pattern beam loop (CTEMECP.f90:292-369): the
    beam sets with the foil normal, DynMat, Sgh, Lgh per thickness in `mode` ('Blochwv' or
    'ScatMat'; exact=True: double-precision q, scipy's expm) and
    sr(t) = fl32(real(sum(Lgh(t) * Sgh)) / float(sum(nat))); returns (sr float32 [nt], the double
    sums, the strong beams, DynMat)."""
    st, ssg, wk, wsg = select_beams_fn(cr, kvec, fn)
    dm = dynmat(cr, st, ssg, wk, wsg)
    S = sgh_ecp(cr, st)
    if mode == "Blochwv":
        L = ecp_lgh_blochwv(dm, kn, thick, round_q=not exact)
    else:
        L = ecp_lgh_scatmat(dm, cr.mlambda, thick, zintstep, exact=exact)
    dots = np.array([np.sum(Lt * S).real for Lt in L])
    sr = (dots / float(np.float32(cr.nat))).astype(np.float32)
    return sr, dots, st, dm


# ----------------------------------------------------------------------------------------
# Monte Carlo electron trajectories (src_programs/CTEMMC.f90)

class _McRng(C.Structure):
    _fields_ = [("s", C.c_int32 * 4)]


def mc_rng_stream(seed, n, jump=0):
    """n draws of rng.f95's generator seeded with `seed` (rng_seed), after a jump of `jump` draws"""
    L = lib()
    L.orc_mcrng_uniform.restype = C.c_float
    r = _McRng()
    L.orc_mcrng_seed(C.byref(r), C.c_int32(seed))
    if jump:
        L.orc_mcrng_jump.argtypes = [C.c_void_p, C.c_uint64]
        L.orc_mcrng_jump(C.byref(r), jump)
    return np.array([L.orc_mcrng_uniform(C.byref(r)) for _ in range(n)], np.float32), tuple(r.s)


def mc_run(p: dict, mode=0, kahan=True):
    """orc_mc_run on the &MCdata values p (sig, omega, numsx, primeseed, num_el, EkeV, Ehistmin,
    Ebinsize, depthmax, depthstep) and CalcDensity's (Ze, density, at_wt, double): accum_e
    [numsx][numsx][numEbins], accum_z [nz][nz][numzbins][numEbins] int32, dropped hits"""
    L = lib()
    D = C.c_double
    L.orc_mc_run.restype = C.c_int64
    L.orc_mc_run.argtypes = [D, D, C.c_int32, C.c_int32, C.c_int64, D, D, D, D, D, D, D, D, C.c_int, C.c_int,
                             C.c_void_p, C.c_void_p]
    nE = int((p["EkeV"] - p["Ehistmin"]) / p["Ebinsize"]) + 1
    nzb = int(p["depthmax"] / p["depthstep"]) + 1
    nx = (p["numsx"] - 1) // 2
    nzw = 2 * (nx // 10) + 1
    ae = np.zeros((p["numsx"], p["numsx"], nE), np.int32)
    az = np.zeros((nzw, nzw, nzb, nE), np.int32)
    dropped = L.orc_mc_run(p["sig"], p["omega"], p["numsx"], p["primeseed"], p["num_el"], p["EkeV"], p["Ehistmin"],
                           p["Ebinsize"], p["depthmax"], p["depthstep"], p["Ze"], p["density"], p["at_wt"], mode,
                           1 if kahan else 0, _p(ae), _p(az))
    return ae, az, int(dropped)
