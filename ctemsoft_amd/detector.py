"""EBSD detector patterns from a master pattern (include/ctem_detector.h) from Python, and a
drop-in for the reference's detector-pattern program CTEMEBSD (src_programs/CTEMEBSD.f90):

    det = detector.Detector(geometry, mc)            # &EBSDdata, the MC energy histograms
    det.set_master(master["sr"])                     # srtmp [numset][nE][2npy+1][2npx+1]
    pats = det.patterns(detector.eu2qu_deg(eulers))  # noise-free float32 [n][numsy][numsx]

    detector.run_program("CTEMEBSD.nml")             # the program: same inputs, same output file

The patterns are computed on the GPU; the noise, binning and gamma are the library's host
routines with the program's semantics. Without the HIP library or a GPU the calls raise
(ebsd.CtemError / RuntimeError); there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import struct
from dataclasses import dataclass

import numpy as np

from . import ebsd
from .masterfile import read_master

EXPORTS = ["ctem_detector_create", "ctem_detector_info", "ctem_detector_set_master", "ctem_detector_patterns",
           "ctem_detector_patterns_device", "ctem_detector_sync", "ctem_detector_last_ms", "ctem_detector_destroy",
           "ctem_detector_last_error", "ctem_eu2qu_deg", "ctem_apply_axisangle", "ctem_poisson_noise",
           "ctem_bin_pattern"]


class _Geometry(C.Structure):
    _fields_ = [("L", C.c_float), ("thetac", C.c_float), ("delta", C.c_float), ("numsx", C.c_int32),
                ("numsy", C.c_int32), ("xpc", C.c_float), ("ypc", C.c_float), ("energymin", C.c_float),
                ("energymax", C.c_float), ("beamcurrent", C.c_double), ("dwelltime", C.c_double)]


class _MC(C.Structure):
    _fields_ = [("numEbins", C.c_int32), ("nsx", C.c_int32), ("nsy", C.c_int32), ("Ehistmin", C.c_double),
                ("Ebinsize", C.c_double), ("MCsig", C.c_double), ("accum_e", C.c_void_p)]


_BOUND = False


def _lib():
    global _BOUND
    L = ebsd.lib()
    if not _BOUND:
        P, I32 = C.c_void_p, C.c_int32
        sig = {"ctem_detector_create": [I32, P, P, P], "ctem_detector_info": [P, P, P, P, P],
               "ctem_detector_set_master": [P, I32, I32, I32, I32, P], "ctem_detector_patterns": [P, I32, P, P],
               "ctem_detector_patterns_device": [P, I32, P, P], "ctem_detector_sync": [P],
               "ctem_detector_last_ms": [P, P], "ctem_detector_destroy": [P], "ctem_eu2qu_deg": [I32, P, P],
               "ctem_apply_axisangle": [I32, P, P], "ctem_poisson_noise": [I32, I32, I32, P, P],
               "ctem_bin_pattern": [I32, I32, I32, I32, C.c_float, P, P]}
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = I32
        L.ctem_detector_last_error.restype = C.c_char_p
        _BOUND = True
    return L


def _check(rc):
    if rc != 0:
        raise ebsd.CtemError(rc, _lib().ctem_detector_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class Geometry:
    """&EBSDdata's detector and beam (CTEMEBSD.f90:167-198; the program's defaults)."""
    L: float = 20000.0
    thetac: float = 0.0
    delta: float = 25.0
    numsx: int = 640
    numsy: int = 480
    xpc: float = 0.0
    ypc: float = 0.0
    energymin: float = 15.0
    energymax: float = 30.0
    beamcurrent: float = 14.513
    dwelltime: float = 100.0


@dataclass
class MCEnergy:
    """The Monte Carlo file's energy histograms: accum_e [2nsy+1][2nsx+1][numEbins] (C order, the
    Fortran accum_e(numEbins, -nsx:nsx, -nsy:nsy)), bins from Ehistmin by Ebinsize [keV], MCsig."""
    accum_e: np.ndarray
    Ehistmin: float
    Ebinsize: float
    MCsig: float

    @property
    def numEbins(self):
        return self.accum_e.shape[2]

    @property
    def nsx(self):
        return (self.accum_e.shape[1] - 1) // 2

    @property
    def nsy(self):
        return (self.accum_e.shape[0] - 1) // 2


class Detector:
    def __init__(self, geometry: Geometry, mc: MCEnergy, device: int = 0):
        L = _lib()
        g = geometry
        self.geometry = g
        self._acc = np.ascontiguousarray(mc.accum_e, np.int32)
        gs = _Geometry(g.L, g.thetac, g.delta, int(g.numsx), int(g.numsy), g.xpc, g.ypc, g.energymin, g.energymax,
                       float(g.beamcurrent), float(g.dwelltime))
        ms = _MC(mc.numEbins, mc.nsx, mc.nsy, float(mc.Ehistmin), float(mc.Ebinsize), float(mc.MCsig),
                 self._acc.ctypes.data)
        h = C.c_void_p()
        _check(L.ctem_detector_create(int(device), C.byref(gs), C.byref(ms), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            _lib().ctem_detector_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        emin, emax, nel = C.c_int32(), C.c_int32(), C.c_int64()
        pf = C.c_double()
        _check(_lib().ctem_detector_info(self.h, C.byref(emin), C.byref(emax), C.byref(nel), C.byref(pf)))
        return dict(Emin=emin.value, Emax=emax.value, num_el=nel.value, prefactor=pf.value)

    def set_master(self, srtmp):
        """srtmp: float32 [numset][nE][2npy+1][2npx+1] (the master file's array)."""
        s = np.ascontiguousarray(srtmp, np.float32)
        nset, nE, ny2, nx2 = s.shape
        _check(_lib().ctem_detector_set_master(self.h, (nx2 - 1) // 2, (ny2 - 1) // 2, nE, nset, _ptr(s)))

    def patterns(self, quats) -> np.ndarray:
        q = np.ascontiguousarray(quats, np.float32).reshape(-1, 4)
        out = np.empty((len(q), self.geometry.numsy, self.geometry.numsx), np.float32)
        _check(_lib().ctem_detector_patterns(self.h, len(q), _ptr(q), _ptr(out)))
        return out

    def patterns_device(self, n: int, d_quats: int, d_patterns: int):
        """device pointers (e.g. torch tensors' data_ptr()); enqueued on the detector's stream"""
        _check(_lib().ctem_detector_patterns_device(self.h, int(n), C.c_void_p(d_quats), C.c_void_p(d_patterns)))

    def sync(self):
        _check(_lib().ctem_detector_sync(self.h))

    def last_ms(self) -> float:
        ms = C.c_double()
        _check(_lib().ctem_detector_last_ms(self.h, C.byref(ms)))
        return ms.value


def eu2qu_deg(eulers) -> np.ndarray:
    e = np.ascontiguousarray(eulers, np.float32).reshape(-1, 3)
    q = np.empty((len(e), 4), np.float32)
    _check(_lib().ctem_eu2qu_deg(len(e), _ptr(e), _ptr(q)))
    return q


def apply_axisangle(quats, axisangle) -> np.ndarray:
    q = np.array(quats, np.float32).reshape(-1, 4)
    a = np.ascontiguousarray(axisangle, np.float32)
    _check(_lib().ctem_apply_axisangle(len(q), _ptr(a), _ptr(q)))
    return q


def new_noise_state(idum: int = -1) -> np.ndarray:
    st = np.zeros(16, np.int64)
    st[0] = idum
    return st


def poisson_noise(patterns, state) -> np.ndarray:
    """POIDEV over the patterns in the program's order, in place on a float32 copy; `state`
    (new_noise_state()) carries the generator from call to call."""
    p = np.array(patterns, np.float32)
    npat = 1 if p.ndim == 2 else p.shape[0]
    _check(_lib().ctem_poisson_noise(p.shape[-1], p.shape[-2], npat, _ptr(p), _ptr(state)))
    return p


def bin_pattern(pattern, binning: int, gamma: bool = False, gammavalue: float = 1.0) -> np.ndarray:
    p = np.ascontiguousarray(pattern, np.float32)
    ny, nx = p.shape
    out = np.empty((ny // binning, nx // binning), np.float32)
    _check(_lib().ctem_bin_pattern(nx, ny, int(binning), int(bool(gamma)), float(gammavalue), _ptr(p), _ptr(out)))
    return out


# ---- the program's files ---------------------------------------------------------------------

def _records(path):
    with open(path, "rb") as f:
        data = f.read()
    recs, off = [], 0
    while off < len(data):
        (n,) = struct.unpack_from("<i", data, off)
        recs.append(data[off + 4: off + 4 + n])
        off += 8 + n
    return recs


def read_mc_energy(path) -> MCEnergy:
    """The Monte Carlo file as CTEMEBSD reads it (CTEMEBSD.f90:299-330; written by
    CTEMMC.f90:264-276): the energy histograms and the sample tilt."""
    r = _records(path)
    numEbins, numzbins, numsx, numsy = struct.unpack_from("<4i", r[3])
    _ekev, ehistmin, ebinsize, _dmax, _dstep = struct.unpack("<5d", r[4])
    mcsig, _omega = struct.unpack("<2d", r[5])
    acc = np.frombuffer(r[7], np.int32).reshape(numsy, numsx, numEbins)
    return MCEnergy(acc.copy(), ehistmin, ebinsize, mcsig)


_NML_DEFAULTS = dict(L=20000.0, thetac=0.0, delta=25.0, numsx=640, numsy=480, xpc=0.0, ypc=0.0,
                     anglefile="euler.txt", eulerconvention="tsl", FZfile="FZ.data", energyfile="energy.data",
                     datafile="EBSDout.data", beamcurrent=14.513, dwelltime=100.0, energymin=15.0,
                     energymax=30.0, binning=1, gammavalue=1.0, scalingmode="not",
                     axisangle=[0.0, 0.0, 0.0, 0.0])


def read_namelist(path) -> dict:
    """&EBSDdata with the program's defaults (CTEMEBSD.f90:167-205)."""
    txt = open(path).read()
    m = re.search(r"&EBSDdata(.*?)/", txt, re.S | re.I)
    if not m:
        raise ValueError(f"{path}: no &EBSDdata group")
    out = dict(_NML_DEFAULTS)
    keys = {k.lower(): k for k in out}
    for name, val in re.findall(r"(\w+)\s*=\s*((?:'[^']*'|\"[^\"]*\"|[^=])*?)(?=,?\s*\w+\s*=|\s*$)", m.group(1), re.S):
        k = keys.get(name.lower())
        if k is None:
            raise ValueError(f"{path}: unknown &EBSDdata entry {name}")
        val = val.strip().rstrip(",").strip()
        if val[:1] in "'\"":
            out[k] = val[1:-1]
        elif k == "axisangle":
            out[k] = [float(v.replace("d", "e").replace("D", "e")) for v in re.split(r"[,\s]+", val) if v]
        elif isinstance(_NML_DEFAULTS[k], int):
            out[k] = int(val)
        else:
            out[k] = float(val.replace("d", "e").replace("D", "e"))
    return out


def read_angles(path):
    """The angle file (CTEMEBSD.f90:215-262): 'eu' or 'qu', the count, then one triplet
    (degrees) or quaternion per line."""
    with open(path) as f:
        kind = f.readline().strip()[:2]
        n = int(f.readline().split()[0])
        vals = np.array([[float(v) for v in f.readline().split()] for _ in range(n)], np.float32)
    return ("euler" if kind == "eu" else "quats"), vals


def run_program(nml_path: str, device: int = 0, chunk: int = 4096, idum: int = -1) -> str:
    """ComputeEBSDPatterns (CTEMEBSD.f90:117-673) end to end: the namelist, angle file, MC file
    and master file in, the program's output file out (record 1: the pattern size and count;
    then one record per pattern, after Poisson noise and, unless scalingmode is 'not', binning
    and gamma). Paths in the namelist are relative to its directory. Returns the output path."""
    base = os.path.dirname(os.path.abspath(nml_path))
    nml = read_namelist(nml_path)
    rel = lambda p: p if os.path.isabs(p) else os.path.join(base, p)
    kind, vals = read_angles(rel(nml["anglefile"]))
    if kind == "euler":
        if nml["eulerconvention"] == "hkl":
            vals[:, 0] = vals[:, 0] + np.float32(90.0)
        quats = eu2qu_deg(vals)
    else:
        quats = vals.reshape(-1, 4)
    quats = apply_axisangle(quats, nml["axisangle"])
    mc = read_mc_energy(rel(nml["energyfile"]))
    master = read_master(rel(nml["FZfile"]))
    geo = Geometry(nml["L"], nml["thetac"], nml["delta"], nml["numsx"], nml["numsy"], nml["xpc"], nml["ypc"],
                   nml["energymin"], nml["energymax"], nml["beamcurrent"], nml["dwelltime"])
    det = Detector(geo, mc, device)
    try:
        det.set_master(master["sr"])
        binning, mode = int(nml["binning"]), nml["scalingmode"]
        out_path = rel(nml["datafile"])
        state = new_noise_state(idum)
        nx, ny = geo.numsx, geo.numsy
        head = (nx, ny) if binning == 1 else (nx // binning, ny // binning)
        with open(out_path, "wb") as f:
            hdr = struct.pack("<3i", head[0], head[1], len(quats))
            f.write(struct.pack("<i", len(hdr)) + hdr + struct.pack("<i", len(hdr)))
            for a0 in range(0, len(quats), chunk):
                pats = det.patterns(quats[a0:a0 + chunk])
                pats = poisson_noise(pats, state)
                for p in pats:
                    if mode != "not":
                        p = bin_pattern(p, binning, mode == "gam", nml["gammavalue"])
                    b = np.ascontiguousarray(p, np.float32).tobytes()
                    f.write(struct.pack("<i", len(b)) + b + struct.pack("<i", len(b)))
        return out_path
    finally:
        det.close()


def read_output(path) -> np.ndarray:
    """The program's output file -> float32 [n][ny][nx]."""
    r = _records(path)
    nx, ny, n = struct.unpack("<3i", r[0])
    return np.stack([np.frombuffer(b, np.float32).reshape(-1, nx) for b in r[1:1 + n]]) if n else \
        np.zeros((0, ny, nx), np.float32)
