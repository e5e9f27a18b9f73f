"""EBSD master-pattern file in the reference's layout (CTEMEBSDmaster.f90:540-568), Fortran
sequential unformatted records with 4-byte markers:

    progname char(132) | scversion char(8) | xtalname char(132) | energyfile char(132)
    npx, npy, numEbins, numset (int32) | EkeVs (float32 x numEbins) | ATOM_type (int32 x numset)
    'square' or 'hexago' char(6) | sr(-npx:npx, -npy:npy, numEbins, numset) float32, Fortran order

The reference samples the northern hemisphere only. When a southern hemisphere is given it is
written as one extra record of the same shape after `sr`; readers of the reference layout
(CTEMEBSD.f90:378-404) stop before it.
"""
from __future__ import annotations

import struct

import numpy as np

SCVERSION = "2.0/2013"


def _rec(b: bytes) -> bytes:
    n = struct.pack("<i", len(b))
    return n + b + n


def _chars(s: str, n: int) -> bytes:
    return s.encode().ljust(n, b" ")[:n]


def write_master(path, sr, ekevs, atom_types, xtalname="crystal.xtal", energyfile="mc.data",
                 progname="CTEMEBSDmaster.f90", grid="square", sr_south=None):
    """sr: float32 array [numset][numEbins][2*npy+1][2*npx+1] (C order, i fastest)."""
    sr = np.asarray(sr, np.float32)
    nset, nE, ny2, nx2 = sr.shape
    npx, npy = (nx2 - 1) // 2, (ny2 - 1) // 2
    recs = [_rec(_chars(progname, 132)), _rec(_chars(SCVERSION, 8)), _rec(_chars(xtalname, 132)),
            _rec(_chars(energyfile, 132)), _rec(struct.pack("<4i", npx, npy, nE, nset)),
            _rec(np.asarray(ekevs, np.float32).tobytes()),
            _rec(np.asarray(atom_types, np.int32).tobytes()),
            _rec(_chars("hexago" if grid == "hexago" else "square", 6)),
            _rec(sr.tobytes())]  # C order [set][E][j][i] == Fortran sr(i,j,E,set)
    if sr_south is not None:
        s = np.asarray(sr_south, np.float32)
        if s.shape != sr.shape:
            raise ValueError("southern hemisphere must have the shape of the northern one")
        recs.append(_rec(s.tobytes()))
    with open(path, "wb") as f:
        f.write(b"".join(recs))


def read_master(path) -> dict:
    with open(path, "rb") as f:
        data = f.read()
    recs, off = [], 0
    while off < len(data):
        (n,) = struct.unpack_from("<i", data, off)
        (n2,) = struct.unpack_from("<i", data, off + 4 + n)
        if n2 != n:
            raise ValueError("corrupt record markers")
        recs.append(data[off + 4: off + 4 + n])
        off += 8 + n
    npx, npy, nE, nset = struct.unpack("<4i", recs[4])
    shape = (nset, nE, 2 * npy + 1, 2 * npx + 1)
    out = dict(progname=recs[0].decode().rstrip(), scversion=recs[1].decode(),
               xtalname=recs[2].decode().rstrip(), energyfile=recs[3].decode().rstrip(),
               npx=npx, npy=npy, nE=nE, nset=nset, ekevs=np.frombuffer(recs[5], np.float32).copy(),
               atom_types=np.frombuffer(recs[6], np.int32).copy(), grid=recs[7].decode(),
               sr=np.frombuffer(recs[8], np.float32).reshape(shape).copy())
    if len(recs) > 9:
        out["sr_south"] = np.frombuffer(recs[9], np.float32).reshape(shape).copy()
    return out
