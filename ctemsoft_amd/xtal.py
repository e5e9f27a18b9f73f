"""Native per-energy setup (include/ctem_setup.h) from Python: the crystal as the .xtal file holds
it, one energy, the Bethe cutoffs -> an ebsd.Setup ready for Context.upload_crystal, plus the
reference's scalar values (wavelength, cell volume, Laue group, ...).

    st, vals = xtal.native_setup(workloads.get("fe_c1").xtal, 20.0, 0.05, (4.0, 8.0, 0.05))
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import ebsd


class _Xtal(C.Structure):
    _fields_ = [("xtal_system", C.c_int32), ("a", C.c_double), ("b", C.c_double), ("c", C.c_double),
                ("alpha", C.c_double), ("beta", C.c_double), ("gamma", C.c_double),
                ("sgnum", C.c_int32), ("sgset", C.c_int32), ("ntype", C.c_int32),
                ("atom_type", C.c_void_p), ("atom_pos", C.c_void_p)]


class _Values(C.Structure):
    _fields_ = [("mlambda", C.c_double), ("mrelcor", C.c_double), ("mpsihat", C.c_double),
                ("msigma", C.c_double), ("vol", C.c_double), ("dynupz", C.c_float),
                ("isym", C.c_int32), ("usehex", C.c_int32), ("nsym", C.c_int32), ("npt", C.c_int32),
                ("dmt", C.c_double * 9), ("rmt", C.c_double * 9), ("dsm", C.c_double * 9),
                ("rsm", C.c_double * 9)]


EXPORTS = ["ctem_setup_create", "ctem_setup_create_gpu", "ctem_setup_create_ecp", "ctem_setup_crystal",
           "ctem_setup_density", "ctem_setup_info", "ctem_setup_destroy", "ctem_setup_last_error"]

_BOUND = False


def _lib():
    global _BOUND
    L = ebsd.lib()
    if not _BOUND:
        P, I32, F = C.c_void_p, C.c_int32, C.c_float
        for name, args in {"ctem_setup_create": [P, F, F, F, F, F, P],
                           "ctem_setup_create_gpu": [I32, P, F, F, F, F, F, P, P], "ctem_setup_crystal": [P, P],
                           "ctem_setup_create_ecp": [P, F, F, P, F, F, F, F, P], "ctem_setup_density": [P, P, P, P],
                           "ctem_setup_info": [P, P], "ctem_setup_destroy": [P]}.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = I32
        L.ctem_setup_last_error.restype = C.c_char_p
        _BOUND = True
    return L


def _check(rc):
    if rc != 0:
        raise ebsd.CtemError(rc, _lib().ctem_setup_last_error().decode())


def native_setup(xtal: dict, ekev: float, dmin: float, cutoffs=(80.0, 160.0, 0.05), device=None, ecp=None):
    """xtal: dict(system, abc, angles, sg, sgset, sites=[dict(Z, xyz, occ, B)]) as in
    ctemsoft_amd.workloads. Returns (ebsd.Setup, values dict). With `device` (a GPU index) the
    look-up-table box is computed on that GPU (ctem_setup_create_gpu) and values["lut_ms"]
    holds the kernel time. ecp = (voltage [V], zone axis k, convergence angle [rad]): the
    channeling-pattern program's set-up (ctem_setup_create_ecp; ekev is then ignored)."""
    L = _lib()
    types = np.array([s["Z"] for s in xtal["sites"]], np.int32)
    pos = np.array([[*s["xyz"], s["occ"], s["B"]] for s in xtal["sites"]], np.float32)
    x = _Xtal(int(xtal["system"]), *[float(v) for v in xtal["abc"]], *[float(v) for v in xtal["angles"]],
              int(xtal["sg"]), int(xtal.get("sgset", 1)), len(types), types.ctypes.data, pos.ctypes.data)
    h = C.c_void_p()
    wc, c, sgc = (float(np.float32(v)) for v in cutoffs)
    ms = C.c_double(-1.0)
    if ecp is not None:
        kz = np.ascontiguousarray(ecp[1], np.int32)
        _check(L.ctem_setup_create_ecp(C.byref(x), float(np.float32(ecp[0])), float(np.float32(dmin)), kz.ctypes.data,
                                       float(np.float32(ecp[2])), wc, c, sgc, C.byref(h)))
    elif device is None:
        _check(L.ctem_setup_create(C.byref(x), float(np.float32(ekev)), float(np.float32(dmin)), wc, c, sgc,
                                   C.byref(h)))
    else:
        _check(L.ctem_setup_create_gpu(int(device), C.byref(x), float(np.float32(ekev)), float(np.float32(dmin)),
                                       wc, c, sgc, C.byref(ms), C.byref(h)))
    try:
        cr = ebsd._Crystal()
        _check(L.ctem_setup_crystal(h, C.byref(cr)))
        v = _Values()
        _check(L.ctem_setup_info(h, C.byref(v)))
        nref = cr.nref
        dims = [cr.lut_dim[i] for i in range(3)]
        ncell = dims[0] * dims[1] * dims[2]
        arr = lambda ptr, ct, n: np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), (n,)).copy()
        hkl = arr(cr.hkl, C.c_int32, 3 * nref).reshape(-1, 3)
        ucg = arr(cr.ucg, C.c_double, 2 * nref).view(np.complex128)
        dbd = arr(cr.dbdiff, C.c_uint8, nref)
        lut = arr(cr.lut, C.c_double, 2 * ncell).view(np.complex128).reshape(dims[2], dims[1], dims[0])
        nsite = cr.nsite
        norb = arr(cr.site_norbit, C.c_int32, nsite)
        orb = arr(cr.orbit, C.c_double, 3 * int(norb.sum())).reshape(-1, 3)
        zz = arr(cr.site_z, C.c_int32, nsite)
        bb = arr(cr.site_b, C.c_float, nsite)
        sites, o = [], 0
        for p in range(nsite):
            sites.append(dict(Z=int(zz[p]), B=float(bb[p]), orbit=orb[o:o + norb[p]]))
            o += int(norb[p])
        vals = dict(mLambda=v.mlambda, mRelcor=v.mrelcor, mPsihat=v.mpsihat, mSigma=v.msigma, vol=v.vol,
                    DynUpz=v.dynupz, isym=v.isym, usehex=v.usehex, nsym=v.nsym, npt=v.npt,
                    dmt=np.array(v.dmt).reshape(3, 3), rmt=np.array(v.rmt).reshape(3, 3),
                    dsm=np.array(v.dsm).reshape(3, 3), rsm=np.array(v.rsm).reshape(3, 3),
                    lut_lo=np.array([cr.lut_lo[i] for i in range(3)], np.int32),
                    lut_ms=ms.value if device is not None and ecp is None else None)
        st = ebsd.Setup(hkl, ucg, dbd, lut, vals["lut_lo"], v.mlambda, v.dynupz, cr.weakcutoff, cr.cutoff,
                        cr.sgcutoff, vals["rmt"], sites, dsm=vals["dsm"], isym=v.isym, usehex=v.usehex)
        return st, vals
    finally:
        L.ctem_setup_destroy(h)


def read_xtal(path: str) -> dict:
    """The .xtal crystal file (files.f90:271-277, two Fortran sequential records: crystal system
    and lattice parameters; ATOM_type(250), ATOM_pos(250,5) float32, ntype, SGnum, SGset) as the
    dict that native_setup takes."""
    import struct
    with open(path, "rb") as f:
        data = f.read()
    recs, off = [], 0
    while off < len(data):
        (n,) = struct.unpack_from("<i", data, off)
        recs.append(data[off + 4: off + 4 + n])
        off += 8 + n
    if len(recs) < 2 or len(recs[0]) != 52:
        raise ValueError(f"{path}: not a .xtal file")
    system, *lat = struct.unpack("<i6d", recs[0])
    types = np.frombuffer(recs[1], np.int32, 250, 0)
    pos = np.frombuffer(recs[1], np.float32, 1250, 1000).reshape(5, 250).T
    ntype, sg, sgset = struct.unpack_from("<3i", recs[1], 6000)
    sites = [dict(Z=int(types[i]), xyz=tuple(float(v) for v in pos[i, :3]), occ=float(pos[i, 3]), B=float(pos[i, 4]))
             for i in range(ntype)]
    return dict(system=int(system), abc=tuple(lat[:3]), angles=tuple(lat[3:]), sg=int(sg), sgset=int(sgset),
                sites=sites)


def density(xtal: dict):
    """CalcDensity (crystal.f90:1388-1417) through the native set-up: (dens, avZ, avA) float32"""
    L = _lib()
    types = np.array([s["Z"] for s in xtal["sites"]], np.int32)
    pos = np.array([[*s["xyz"], s["occ"], s["B"]] for s in xtal["sites"]], np.float32)
    x = _Xtal(int(xtal["system"]), *[float(v) for v in xtal["abc"]], *[float(v) for v in xtal["angles"]],
              int(xtal["sg"]), int(xtal.get("sgset", 1)), len(types), types.ctypes.data, pos.ctypes.data)
    h = C.c_void_p()
    # the density needs the cell and the orbits only: a coarse dmin keeps the reflection list tiny
    _check(L.ctem_setup_create(C.byref(x), 20.0, 10.0, 4.0, 8.0, 0.05, C.byref(h)))
    try:
        d, z, a = C.c_float(), C.c_float(), C.c_float()
        _check(L.ctem_setup_density(h, C.byref(d), C.byref(z), C.byref(a)))
        return np.float32(d.value), np.float32(z.value), np.float32(a.value)
    finally:
        L.ctem_setup_destroy(h)
