"""Electron channeling patterns (include/ctem_ecp.h) from Python.

The reference program ECpattern (src_programs/CTEMECP.f90:70-452) as a function: its set-up
through the native library (crystal, wavelength from the voltage in V, foil normal, ga,
convergence, reflection list with the convergence-angle window, incident directions), the beam
loop on the GPU in either of the program's solution modes, and the result as the program's
sr(2 npix + 1, 2 npix + 1, numthick) array.

    from ctemsoft_amd import ecp, workloads
    r = ecp.compute(workloads.get("ni_c0").xtal, voltage=20000.0, k=(0, 0, 1), fn=(0, 0, 1), dmin=0.05,
                    thetac=120.0, startthick=2.0, thickinc=2.0, numthick=10, npix=64, compmode="Blochwv")
    r["sr"]            # float32 [numthick][2 npix + 1][2 npix + 1]

Errors raise ebsd.CtemError (the library's status code and message); nothing runs without the
HIP library and a GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import ebsd, xtal as xtal_mod

MODES = {"Blochwv": 0, "ScatMat": 1}

EXPORTS = ["ctem_ecp_last_error", "ctem_ecp_foil_normal", "ctem_ecp_shortest_g", "ctem_ecp_convergence",
           "ctem_ecp_kvectors", "ctem_ecp_plan_create", "ctem_ecp_run", "ctem_ecp_get_sr", "ctem_ecp_image"]


class _Params(C.Structure):
    _fields_ = [("mode", C.c_int32), ("numthick", C.c_int32), ("startthick", C.c_float),
                ("thickinc", C.c_float), ("zintstep", C.c_float), ("fn", C.c_float * 3),
                ("site_occ", C.c_void_p)]


_BOUND = False


def _lib():
    global _BOUND
    L = ebsd.lib()
    if not _BOUND:
        P, I32, F, D = C.c_void_p, C.c_int32, C.c_float, C.c_double
        for name, args in {"ctem_ecp_foil_normal": [P, P, P, P], "ctem_ecp_shortest_g": [P, P, P, P],
                           "ctem_ecp_convergence": [P, P, D, P, P, P, P],
                           "ctem_ecp_kvectors": [P, P, F, I32, P, P, D, I32, P, P, P, P, P],
                           "ctem_ecp_plan_create": [P, I32, P, P, P, P], "ctem_ecp_run": [P],
                           "ctem_ecp_get_sr": [P, P], "ctem_ecp_image": [P, P, I32, P]}.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = I32
        L.ctem_ecp_last_error.restype = C.c_char_p
        _BOUND = True
    return L


def _gcheck(rc):
    if rc != 0:
        raise ebsd.CtemError(rc, _lib().ctem_ecp_last_error().decode())


def _m9(m):
    return np.ascontiguousarray(np.asarray(m, np.float64).reshape(3, 3)).ravel()


def foil_normal(fn, dmt, rmt):
    """DynFN of the program (CTEMECP.f90:204-205), float32 reciprocal components."""
    out = np.zeros(3, np.float32)
    f = np.ascontiguousarray(fn, np.int32)
    _gcheck(_lib().ctem_ecp_foil_normal(ebsd._p(f), ebsd._p(_m9(dmt)), ebsd._p(_m9(rmt)), ebsd._p(out)))
    return out


def shortest_g(k, dmt, rmt):
    """ga of ShortestG (symmetry.f90:1133-1309)."""
    out = np.zeros(3, np.int32)
    kk = np.ascontiguousarray(k, np.int32)
    _gcheck(_lib().ctem_ecp_shortest_g(ebsd._p(kk), ebsd._p(_m9(dmt)), ebsd._p(_m9(rmt)), ebsd._p(out)))
    return out


def convergence(ga, rmt, mlambda, thetac=0.0, ktmax=0.0):
    """(thetac [mrad], ktmax, bragg, galen) of CTEMECP.f90:227-239: ktmax from thetac unless
    ktmax != 0."""
    g = np.ascontiguousarray(ga, np.int32)
    th, kt, br, gl = C.c_float(thetac), C.c_float(ktmax), C.c_float(), C.c_float()
    _gcheck(_lib().ctem_ecp_convergence(ebsd._p(g), ebsd._p(_m9(rmt)), float(mlambda), C.byref(th), C.byref(kt),
                                        C.byref(br), C.byref(gl)))
    return th.value, kt.value, br.value, gl.value


def kvectors(k, ga, ktmax, npx, dmt, rmt, mlambda):
    """CalckvectorsSymmetry of the program (kvectors.f90:573-690): dict(ij, k, kt, kn)."""
    L = _lib()
    kk = np.ascontiguousarray(k, np.int32)
    g = np.ascontiguousarray(ga, np.int32)
    d9, r9 = _m9(dmt), _m9(rmt)
    n = C.c_int32()
    _gcheck(L.ctem_ecp_kvectors(ebsd._p(kk), ebsd._p(g), float(np.float32(ktmax)), int(npx), ebsd._p(d9), ebsd._p(r9),
                                float(mlambda), 0, None, None, None, None, C.byref(n)))
    m = n.value
    ij = np.zeros((m, 2), np.int32)
    kv = np.zeros((m, 3))
    kt = np.zeros((m, 3))
    kn = np.zeros(m)
    _gcheck(L.ctem_ecp_kvectors(ebsd._p(kk), ebsd._p(g), float(np.float32(ktmax)), int(npx), ebsd._p(d9), ebsd._p(r9),
                                float(mlambda), m, ebsd._p(ij), ebsd._p(kv), ebsd._p(kt), ebsd._p(kn), C.byref(n)))
    return dict(ij=ij, k=kv, kt=kt, kn=kn)


class EcpPlan(ebsd.Plan):
    """A channeling-pattern plan: ctem_ecp_plan_create over the directions k, kn on the context's
    current crystal; select / beams / dynmat / timing as for a master-pattern plan."""

    def __init__(self, ctx: ebsd.Context, k, kn, compmode, startthick, thickinc, numthick, zintstep, dynfn,
                 site_occ):
        self.ctx = ctx
        self.k = np.ascontiguousarray(k, np.float64).reshape(-1, 3)
        self.kn = np.ascontiguousarray(kn, np.float64)
        self.numk = len(self.k)
        self.numthick = int(numthick)
        self.setup = ctx.setup
        self.nsite = len(ctx.setup.sites) if ctx.setup is not None else 0
        self.occ = np.ascontiguousarray(site_occ, np.float32)
        prm = _Params(MODES[compmode], self.numthick, float(np.float32(startthick)), float(np.float32(thickinc)),
                      float(np.float32(zintstep)), (C.c_float * 3)(*[float(v) for v in np.asarray(dynfn, np.float32)]),
                      self.occ.ctypes.data)
        self._h = C.c_void_p()
        ebsd._check(_lib().ctem_ecp_plan_create(ctx._h, self.numk, ebsd._p(self.k), ebsd._p(self.kn), C.byref(prm),
                                                C.byref(self._h)))
        ctx._plans.add(self)
        self.nns = self.nnw = None

    def run(self):
        ebsd._check(_lib().ctem_ecp_run(self._h))

    def sr(self):
        """float32 [numthick][numk]"""
        out = np.zeros((self.numthick, self.numk), np.float32)
        ebsd._check(_lib().ctem_ecp_get_sr(self._h, ebsd._p(out)))
        return out

    def image(self, ij, npx):
        """float32 [numthick][2 npx + 1][2 npx + 1], the program's sr(ipx, ipy, it) as [it][ipy][ipx]"""
        w = 2 * int(npx) + 1
        out = np.zeros((self.numthick, w, w), np.float32)
        ijc = np.ascontiguousarray(ij, np.int32)
        ebsd._check(_lib().ctem_ecp_image(self._h, ebsd._p(ijc), int(npx), ebsd._p(out)))
        return out


def compute(xtal: dict, voltage: float, k, fn, dmin: float, thetac: float = 0.0, startthick: float = 2.0,
            thickinc: float = 2.0, numthick: int = 10, npix: int = 256, compmode: str = "Blochwv",
            zintstep: float = 1.0, cutoffs=(80.0, 160.0, 0.05), ktmax: float = 0.0, device: int = 0, ctx=None):
    """ECpattern (CTEMECP.f90:70-380) with the &ECPlist values (defaults :136-151, except npix and
    the Bethe cutoffs, which the program reads from BetheParameters.nml) on GPU `device` (or the
    given ebsd.Context). Returns dict(sr [numthick][2 npix + 1][2 npix + 1], numk, ga, thetac,
    ktmax, galen, delta, dynfn, plan timing)."""
    if compmode not in MODES:
        raise ValueError(f"compmode must be one of {sorted(MODES)}")
    k = [int(v) for v in k]
    fn = [int(v) for v in fn]
    # crystal metric and wavelength (a set-up without the window gives the same metric)
    st0, v0 = xtal_mod.native_setup(xtal, float(np.float32(voltage)) / 1000.0, dmin, cutoffs,
                                    ecp=(voltage, k, 0.0))
    dmt, rmt, mlambda = v0["dmt"], v0["rmt"], v0["mLambda"]
    dynfn = foil_normal(fn, dmt, rmt)
    ga = shortest_g(k, dmt, rmt)
    thetac, ktmax, bragg, galen = convergence(ga, rmt, mlambda, thetac, ktmax)
    st, vals = xtal_mod.native_setup(xtal, float(np.float32(voltage)) / 1000.0, dmin, cutoffs,
                                     ecp=(voltage, k, np.float32(thetac) / np.float32(1000.0)))
    kv = kvectors(k, ga, ktmax, npix, dmt, rmt, mlambda)
    own = ctx is None
    if own:
        ctx = ebsd.Context(device)
    try:
        ctx.upload_crystal(st)
        occ = np.array([s["occ"] for s in xtal["sites"]], np.float32)
        plan = EcpPlan(ctx, kv["k"], kv["kn"], compmode, startthick, thickinc, numthick, zintstep, dynfn, occ)
        try:
            plan.run()
            img = plan.image(kv["ij"], npix)
            tm = plan.timing()
        finally:
            plan.close()
    finally:
        if own:
            ctx.close()
    delta = np.float32(2.0) * np.float32(ktmax) * np.float32(galen) / np.float32(2 * npix + 1)
    return dict(sr=img, numk=len(kv["k"]), ga=ga, thetac=thetac, ktmax=ktmax, galen=galen, bragg=bragg,
                delta=float(delta), dynfn=dynfn, ij=kv["ij"], timing=tm, nref=len(st.hkl))
