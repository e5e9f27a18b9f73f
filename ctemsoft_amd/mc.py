"""Monte Carlo electron trajectories (include/ctem_mc.h) from Python.

The reference program CTEMMC (src_programs/CTEMMC.f90) as a function and as a drop-in runner on
its own files: the &MCdata namelist and the .xtal crystal in, the program's output file out (the
backscattered-electron energy histograms on the square Lambert grid and the depth histograms).

    from ctemsoft_amd import mc, workloads
    h = mc.simulate(workloads.get("ni_c0").xtal, num_el=10_000_000, EkeV=20.0, Ehistmin=10.0, Ebinsize=1.0)
    h["accum_e"]          # int32 [numsx][numsx][numEbins]
    mc.run_program("CTEMMC.nml")          # writes the program's output file

Modes (ctem_mc.h): "blocked" (default; every electron an own block of the program's random stream,
electrons in parallel) and "sequential" (the program's one stream in its order: its exact output,
one GPU lane). Errors raise ebsd.CtemError; nothing runs without the HIP library and a GPU.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import struct

import numpy as np

from . import ebsd, xtal as xtal_mod

MODES = {"blocked": 0, "sequential": 1}

EXPORTS = ["ctem_mc_create", "ctem_mc_dims", "ctem_mc_run", "ctem_mc_reset", "ctem_mc_get", "ctem_mc_device",
           "ctem_mc_steps", "ctem_mc_last_ms", "ctem_mc_destroy", "ctem_mc_last_error"]

# &MCdata defaults (CTEMMC.f90:153-166)
NML_DEFAULTS = dict(xtalname="undefined", sig=70.0, omega=0.0, numsx=1501, primeseed=932117, num_el=10000000,
                    nthreads=4, EkeV=30.0, Ehistmin=20.0, Ebinsize=0.25, depthmax=50.0, depthstep=1.0,
                    dataname="undefined.data", MCmode="CSDA")


class _Params(C.Structure):
    _fields_ = [("sig", C.c_double), ("omega", C.c_double), ("numsx", C.c_int32), ("primeseed", C.c_int32),
                ("num_el", C.c_int64), ("EkeV", C.c_double), ("Ehistmin", C.c_double), ("Ebinsize", C.c_double),
                ("depthmax", C.c_double), ("depthstep", C.c_double), ("Ze", C.c_double), ("density", C.c_double),
                ("at_wt", C.c_double)]


_BOUND = False


def _lib():
    global _BOUND
    L = ebsd.lib()
    if not _BOUND:
        P, I32, I64 = C.c_void_p, C.c_int32, C.c_int64
        for name, args in {"ctem_mc_create": [I32, P, P], "ctem_mc_dims": [P, P, P, P, P],
                           "ctem_mc_run": [P, I32, I64, I64], "ctem_mc_reset": [P], "ctem_mc_get": [P, P, P, P],
                           "ctem_mc_device": [P, P, P], "ctem_mc_steps": [P, P], "ctem_mc_last_ms": [P, P],
                           "ctem_mc_destroy": [P]}.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = I32
        L.ctem_mc_last_error.restype = C.c_char_p
        _BOUND = True
    return L


def _check(rc):
    if rc != 0:
        raise ebsd.CtemError(rc, _lib().ctem_mc_last_error().decode())


class MonteCarlo:
    """ctem_mc object: derived constants and the histograms resident on `device`."""

    def __init__(self, sig, omega, numsx, primeseed, num_el, EkeV, Ehistmin, Ebinsize, depthmax, depthstep, Ze,
                 density, at_wt, device=0):
        self.p = _Params(float(sig), float(omega), int(numsx), int(primeseed), int(num_el), float(EkeV),
                         float(Ehistmin), float(Ebinsize), float(depthmax), float(depthstep), float(Ze),
                         float(density), float(at_wt))
        self._h = C.c_void_p()
        _check(_lib().ctem_mc_create(int(device), C.byref(self.p), C.byref(self._h)))
        ne, nz, nx, nzw = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        _check(_lib().ctem_mc_dims(self._h, C.byref(ne), C.byref(nz), C.byref(nx), C.byref(nzw)))
        self.numEbins, self.numzbins, self.nx, self.nzw = ne.value, nz.value, nx.value, nzw.value

    def run(self, mode="blocked", first=0, count=None):
        count = self.p.num_el if count is None else count
        _check(_lib().ctem_mc_run(self._h, MODES[mode], int(first), int(count)))

    def reset(self):
        _check(_lib().ctem_mc_reset(self._h))

    def histograms(self):
        n = self.p.numsx
        ae = np.zeros((n, n, self.numEbins), np.int32)
        az = np.zeros((self.nzw, self.nzw, self.numzbins, self.numEbins), np.int32)
        dr = C.c_int64()
        _check(_lib().ctem_mc_get(self._h, ebsd._p(ae), ebsd._p(az), C.byref(dr)))
        return ae, az, dr.value

    def steps(self):
        """trajectory steps since the last reset"""
        n = C.c_int64()
        _check(_lib().ctem_mc_steps(self._h, C.byref(n)))
        return n.value

    def last_ms(self):
        ms = C.c_double()
        _check(_lib().ctem_mc_last_ms(self._h, C.byref(ms)))
        return ms.value

    def close(self):
        if self._h:
            _lib().ctem_mc_destroy(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def simulate(xtal: dict, mode="blocked", device=0, **nml):
    """DoMCsimulation (CTEMMC.f90:103-277) without the file: CalcDensity on the crystal, then num_el
    electrons. nml: the &MCdata entries (defaults NML_DEFAULTS). Returns dict(accum_e, accum_z,
    dropped, ms, dens, avZ, avA, numEbins, numzbins)."""
    p = dict(NML_DEFAULTS)
    p.update(nml)
    dens, avz, ava = xtal_mod.density(xtal)
    m = MonteCarlo(p["sig"], p["omega"], p["numsx"], p["primeseed"], p["num_el"], p["EkeV"], p["Ehistmin"],
                   p["Ebinsize"], p["depthmax"], p["depthstep"], float(avz), float(dens), float(ava), device=device)
    try:
        m.run(mode)
        ae, az, dr = m.histograms()
        return dict(accum_e=ae, accum_z=az, dropped=dr, ms=m.last_ms(), dens=dens, avZ=avz, avA=ava,
                    numEbins=m.numEbins, numzbins=m.numzbins)
    finally:
        m.close()


def read_namelist(path) -> dict:
    """&MCdata with the program's defaults (CTEMMC.f90:150-171)"""
    txt = open(path).read()
    m = re.search(r"&MCdata(.*?)/", txt, re.S | re.I)
    if not m:
        raise ValueError(f"{path}: no &MCdata group")
    out = dict(NML_DEFAULTS)
    keys = {k.lower(): k for k in out}
    for name, val in re.findall(r"(\w+)\s*=\s*((?:'[^']*'|\"[^\"]*\"|[^=])*?)(?=,?\s*\w+\s*=|\s*$)", m.group(1), re.S):
        k = keys.get(name.lower())
        if k is None:
            raise ValueError(f"{path}: unknown &MCdata entry {name}")
        val = val.strip().rstrip(",").strip()
        if val[:1] in "'\"":
            out[k] = val[1:-1]
        elif isinstance(NML_DEFAULTS[k], int):
            out[k] = int(val)
        else:
            out[k] = float(val.replace("d", "e").replace("D", "e"))
    return out


def _rec(b: bytes) -> bytes:
    n = struct.pack("<i", len(b))
    return n + b + n


def write_output(path, nml: dict, accum_e, accum_z):
    """The program's output file (CTEMMC.f90:264-276): program name, version, crystal name, the
    dimensions and electron count (num_el times its one loop pass), energies and depths, tilts,
    mode, then accum_e and accum_z in Fortran order."""
    ae = np.ascontiguousarray(accum_e, np.int32)
    az = np.ascontiguousarray(accum_z, np.int32)
    numEbins, numzbins = ae.shape[2], az.shape[2]
    recs = [
        "CTEMMC.f90".ljust(132).encode(),
        b"2.0/2013",
        str(nml["xtalname"]).ljust(132)[:132].encode(),
        struct.pack("<4iqi", numEbins, numzbins, nml["numsx"], nml["numsx"], int(nml["num_el"]), 1),
        struct.pack("<5d", nml["EkeV"], nml["Ehistmin"], nml["Ebinsize"], nml["depthmax"], nml["depthstep"]),
        struct.pack("<2d", nml["sig"], nml["omega"]),
        str(nml["MCmode"]).ljust(4)[:4].encode(),
        ae.tobytes(),
        az.tobytes(),
    ]
    with open(path, "wb") as f:
        f.write(b"".join(_rec(r) for r in recs))


def run_program(nml_path, mode="sequential", device=0):
    """Drop-in for the program: reads the namelist and the crystal file it names (relative to the
    namelist's folder), runs the electrons and writes the output file `dataname`. The program runs
    its electrons on one random stream (`nthreads` does not split them: its thread loop runs once,
    CTEMMC.f90:222-229); mode "sequential" reproduces that stream, "blocked" runs the electrons in
    parallel on blocks of it. Returns the output file's path."""
    base = os.path.dirname(os.path.abspath(nml_path))
    nml = read_namelist(nml_path)
    if nml["xtalname"] == "undefined":
        raise ValueError("CTEMMC: structure file name is undefined")
    x = xtal_mod.read_xtal(os.path.join(base, nml["xtalname"]))
    r = simulate(x, mode=mode, device=device, **{k: v for k, v in nml.items()
                                                  if k not in ("xtalname", "dataname", "nthreads", "MCmode")})
    out = os.path.join(base, nml["dataname"])
    write_output(out, nml, r["accum_e"], r["accum_z"])
    return out
