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
