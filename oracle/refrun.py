"""Run the compiled reference (oracle/_ref/, built by `make -C oracle ref`) on inputs written
here: the .xtal crystal file, the Monte Carlo depth file, the namelists, and the
instrumentation driver's stream output.

TEST INFRASTRUCTURE ONLY: used by tests/golden/make_golden.py to write fixtures and by
bench.py's CPU-baseline leg to time the reference's own beam loop. Never used by the
product.

File layouts written and read here follow the reference:
  .xtal       files.f90:271-277 (two Fortran sequential records)
  MC file     CTEMMC.f90:264-276 / CTEMEBSDmaster.f90:190-235
  master file CTEMEBSDmaster.f90:540-568
"""
from __future__ import annotations

import os
import struct
import subprocess
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

REF_BIN = os.path.join(HERE, "_ref")
MAXPASYM = 250


# ----------------------------------------------------------------------------------------
# Fortran sequential unformatted records (4-byte little-endian markers)
def _rec(buf: bytes) -> bytes:
    n = struct.pack("<i", len(buf))
    return n + buf + n


def _chars(s: str, n: int) -> bytes:
    return s.encode().ljust(n, b" ")[:n]


def write_xtal(path: str, xtal: dict) -> None:
    """.xtal writer (files.f90:271-277): R1 = system, a,b,c,alpha,beta,gamma;
    R2 = ATOM_type(250), ATOM_pos(250,5) f32 column-major, ntype, SGnum, SGset."""
    r1 = struct.pack("<i6d", xtal["system"], *xtal["abc"], *xtal["angles"])
    types = np.zeros(MAXPASYM, np.int32)
    pos = np.zeros((MAXPASYM, 5), np.float32)
    for i, s in enumerate(xtal["sites"]):
        types[i] = s["Z"]
        pos[i, :3] = s["xyz"]
        pos[i, 3] = s["occ"]
        pos[i, 4] = s["B"]
    r2 = types.tobytes() + pos.T.copy().tobytes() + struct.pack(
        "<3i", len(xtal["sites"]), xtal["sg"], xtal.get("sgset", 1))
    with open(path, "wb") as f:
        f.write(_rec(r1) + _rec(r2))


def write_mcfile(path: str, xtalname: str, ekev: float, dz: float, counts, tail) -> None:
    Z = len(counts)
    nzb = Z + 2
    acc_z = np.zeros(nzb, np.int32)
    acc_z[:Z] = counts
    acc_z[Z] = tail
    num_el = int(np.sum(counts))
    recs = [
        _rec(_chars("CTEMMC.f90", 132)),
        _rec(_chars("2.0/2013", 8)),
        _rec(_chars(xtalname, 132)),
        _rec(struct.pack("<4iqi", 1, nzb, 1, 1, num_el, 1)),
        _rec(struct.pack("<5d", ekev, ekev, 1.0, dz * nzb, dz)),
        _rec(struct.pack("<2d", 70.0, 0.0)),
        _rec(_chars("full", 4)),
        _rec(np.array([num_el], np.int32).tobytes()),
        _rec(acc_z.tobytes()),
    ]
    with open(path, "wb") as f:
        f.write(b"".join(recs))


def read_master(path: str) -> dict:
    with open(path, "rb") as f:
        data = f.read()
    recs, off = [], 0
    while off < len(data):
        (n,) = struct.unpack_from("<i", data, off)
        recs.append(data[off + 4: off + 4 + n])
        off += 8 + n
    npx, npy, nE, nset = struct.unpack("<4i", recs[4])
    sr = np.frombuffer(recs[8], np.float32).reshape((nset, nE, 2 * npy + 1, 2 * npx + 1))
    return dict(npx=npx, npy=npy, nE=nE, nset=nset,
                ekevs=np.frombuffer(recs[5], np.float32).copy(),
                atom_types=np.frombuffer(recs[6], np.int32).copy(),
                grid=recs[7].decode(), sr=sr.copy())


# ----------------------------------------------------------------------------------------
class Stream:
    def __init__(self, buf: bytes):
        self.b, self.o = buf, 0

    def get(self, dtype, count):
        dt = np.dtype(dtype)
        a = np.frombuffer(self.b, dt, count, self.o)
        self.o += dt.itemsize * count
        return a.copy()


def parse_dump(path: str, dolgh: int, lutfull: bool = False) -> dict:
    with open(path, "rb") as f:
        s = Stream(f.read())
    magic, isym, usehex, numk, nref, nsel, nsite, sgnum = s.get("<i4", 8)
    assert magic == 20260
    dims = s.get("<i4", 6)
    consts = s.get("<f8", 9)
    mats = s.get("<f8", 36).reshape(4, 3, 3)  # Fortran column-major 3x3 each
    mats = np.transpose(mats, (0, 2, 1))
    nlut = int(dims[0] * dims[1] * dims[2])
    lut = s.get("<c16", nlut).reshape(dims[2], dims[1], dims[0])  # [l][k][h]
    dbd = s.get("<i4", nlut).reshape(dims[2], dims[1], dims[0])
    hkl = np.zeros((nref, 3), np.int32)
    ucg = np.zeros(nref, np.complex128)
    for i in range(nref):
        hkl[i] = s.get("<i4", 3)
        ucg[i] = s.get("<c16", 1)[0]
    sites = []
    for _ in range(nsite):
        Z = int(s.get("<i4", 1)[0])
        pos = s.get("<f4", 5)
        n = int(s.get("<i4", 1)[0])
        orb = s.get("<f8", 3 * n).reshape(3, n).T.copy()
        sites.append(dict(Z=Z, pos=pos, orbit=orb))
    kij = np.zeros((numk, 2), np.int32)
    kvec = np.zeros((numk, 3), np.float64)
    kn = np.zeros(numk, np.float64)
    for i in range(numk):
        kij[i] = s.get("<i4", 2)
        kvec[i] = s.get("<f8", 3)
        kn[i] = s.get("<f8", 1)[0]
    full = {}
    if lutfull:
        hm = s.get("<i4", 3)
        fd = 4 * hm + 1
        full = dict(lut_full=s.get("<c16", int(np.prod(fd))).reshape(fd[2], fd[1], fd[0]),
                    lut_full_lo=(-2 * hm).astype(np.int32))
    hkl_index = {tuple(h): i for i, h in enumerate(hkl)}
    per = []
    for _ in range(nsel):
        ik, nns, nnw = s.get("<i4", 3)
        strong = s.get("<i4", nns) - 1
        ssg = s.get("<f4", nns)
        whkl = s.get("<i4", 3 * nnw).reshape(nnw, 3)
        wsg = s.get("<f4", nnw)
        weak = np.array([hkl_index[tuple(h)] for h in whkl], np.int32)
        d = dict(ik=int(ik) - 1, strong=strong, strongsg=ssg, weak=weak, weaksg=wsg)
        if dolgh >= 1:
            d["dynmat"] = s.get("<c16", nns * nns).reshape(nns, nns).T.copy()
        if dolgh >= 2:
            d["lgh"] = s.get("<c16", nns * nns).reshape(nns, nns).T.copy()
        per.append(d)
    assert s.o == len(s.b), "trailing bytes in dump"
    return dict(isym=int(isym), usehex=int(usehex), numk=int(numk), nref=int(nref),
                sgnum=int(sgnum), lut_dims=dims, mLambda=consts[0], mRelcor=consts[1],
                mPsihat=consts[2], mSigma=consts[3], DynUpz=consts[4], weakcutoff=consts[5],
                cutoff=consts[6], sgcutoff=consts[7], vol=consts[8], dmt=mats[0], rmt=mats[1],
                dsm=mats[2], rsm=mats[3], lut=lut, dbdiff=dbd, hkl=hkl, ucg=ucg, sites=sites,
                kij=kij, kvec=kvec, kn=kn, per=per, **full)


def run_driver(work: str, wl: workloads.Workload, lam: np.ndarray, ksel, dolgh: int,
               ekev: float | None = None, lutfull: bool = False, kvec=None) -> dict:
    """The reference's beam loop body (ref_driver.f90) on the directions ksel (indices of its own
    list, ascending), or, with kvec = (k [n][3], kn [n]), on those wave vectors instead (e.g. the
    library's southern hemisphere, which the program does not compute); ik then counts them."""
    xt = os.path.join(work, "crystal.xtal")
    write_xtal(xt, wl.xtal)
    with open(os.path.join(work, "BetheParameters.nml"), "w") as f:
        f.write(f" &BetheList\n weakcutoff = {wl.weakcutoff!r},\n cutoff = {wl.cutoff!r},\n"
                f" sgcutoff = {wl.sgcutoff!r},\n /\n")
    lam.astype(np.float32).tofile(os.path.join(work, "lam.bin"))
    ksf = "undefined"
    if ksel is not None:
        ksf = "ksel.bin"
        ks = np.asarray(ksel, np.int32) + 1
        with open(os.path.join(work, ksf), "wb") as f:
            f.write(struct.pack("<i", len(ks)) + ks.tobytes())
    kf = "undefined"
    if kvec is not None:
        kf = "kvec.bin"
        kk, kn = np.asarray(kvec[0], np.float64), np.asarray(kvec[1], np.float64)
        with open(os.path.join(work, kf), "wb") as f:
            f.write(struct.pack("<i", len(kn)) + np.concatenate([kk, kn[:, None]], axis=1).tobytes())
    ek = wl.ekev if ekev is None else ekev
    dz = wl.dz_for(ekev)
    thick = float(int(len(lam) * dz))  # integer thickness (CTEMEBSDmaster.f90:142,306)
    with open(os.path.join(work, "refdrv.nml"), "w") as f:
        f.write(f" &refdrv\n xtalname = 'crystal.xtal',\n ekev = {ek!r},\n dmin = {wl.dmin!r},\n"
                f" npx = {wl.npx},\n izz = {len(lam)},\n depthstep = {dz!r},\n"
                f" thick = {thick!r},\n lamfile = 'lam.bin',\n kselfile = '{ksf}',\n"
                f" outname = 'dump.bin',\n dolgh = {dolgh},\n lutfull = {int(lutfull)},\n kfile = '{kf}',\n /\n")
    env = dict(os.environ, MKL_NUM_THREADS="1", OMP_NUM_THREADS="1")
    subprocess.run(["bash", "-c", f"ulimit -s unlimited && {REF_BIN}/ctemref_driver > driver.log 2>&1"],
                   cwd=work, check=True, env=env)
    d = parse_dump(os.path.join(work, "dump.bin"), dolgh, lutfull)
    d["thick"] = thick
    return d


def run_program(work: str, wl: workloads.Workload, counts, tail) -> dict:
    write_xtal(os.path.join(work, "crystal.xtal"), wl.xtal)
    write_mcfile(os.path.join(work, "mc.data"), "crystal.xtal", wl.ekev, wl.dz, counts, tail)
    with open(os.path.join(work, "BetheParameters.nml"), "w") as f:
        f.write(f" &BetheList\n weakcutoff = {wl.weakcutoff!r},\n cutoff = {wl.cutoff!r},\n"
                f" sgcutoff = {wl.sgcutoff!r},\n /\n")
    with open(os.path.join(work, "CTEMEBSDmaster.nml"), "w") as f:
        f.write(f" &EBSDmastervars\n dmin = {wl.dmin!r},\n npx = {wl.npx},\n"
                f" outname = 'master.data',\n energyfile = 'mc.data',\n Esel = -1,\n /\n")
    env = dict(os.environ, MKL_NUM_THREADS="1", OMP_NUM_THREADS="1")
    subprocess.run(["bash", "-c", f"ulimit -s unlimited && {REF_BIN}/CTEMEBSDmaster CTEMEBSDmaster.nml "
                    "> program.log 2>&1"], cwd=work, check=True, env=env)
    return read_master(os.path.join(work, "master.data"))


def run_sym(work: str, nbox: int, npx: int, dsm: np.ndarray, wavelength: float) -> dict:
    """Laue-equivalence table over the box -nbox..nbox squared (isym 1..13) and the
    RoscaLambert direction lists of every Laue group on both grids, from the reference's
    Apply2DLaueSymmetry and Calckvectors (oracle/ref_sym.f90). dsm is the direct structure
    matrix as a row-major 3x3 array."""
    d = np.asarray(dsm, np.float64)
    vals = ", ".join(repr(float(v)) for v in d.T.reshape(-1))  # Fortran column order
    with open(os.path.join(work, "refsym.nml"), "w") as f:
        f.write(f" &refsym\n nbox = {nbox},\n npx = {npx},\n dsm = {vals},\n"
                f" wavelength = {float(wavelength)!r},\n outname = 'sym.bin',\n /\n")
    subprocess.run(["bash", "-c", f"ulimit -s unlimited && {REF_BIN}/ctemref_sym > sym.log 2>&1"],
                   cwd=work, check=True)
    with open(os.path.join(work, "sym.bin"), "rb") as f:
        s = Stream(f.read())
    magic, nb, npx_ = s.get("<i4", 3)
    assert magic == 20261 and nb == nbox and npx_ == npx
    w = 2 * nbox + 1
    raw = s.get("<i4", 13 * w * w * 25).reshape(13, w, w, 25)  # [isym-1][ipy][ipx]
    out = dict(laue_n=raw[..., 0].astype(np.int8), laue_q=raw[..., 1:].astype(np.int8))
    assert np.array_equal(out["laue_q"].astype(np.int32), raw[..., 1:])
    for _ in range(24):
        isym, ihex, numk = (int(v) for v in s.get("<i4", 3))
        rec = np.frombuffer(s.b, np.dtype([("i", "<i4"), ("j", "<i4"), ("k", "<f8", 3), ("kn", "<f8")]),
                            numk, s.o)
        s.o += rec.itemsize * numk
        out[f"kv{isym}_{ihex}_ij"] = np.stack([rec["i"], rec["j"]], 1).astype(np.int32)
        out[f"kv{isym}_{ihex}_k"] = rec["k"].copy()
        out[f"kv{isym}_{ihex}_kn"] = rec["kn"].copy()
    lg = s.get("<i4", 3 * 230).reshape(230, 3)
    assert np.array_equal(lg[:, 0], np.arange(1, 231))
    out["laue_isym"] = lg[:, 1].astype(np.int8)
    out["laue_usehex"] = lg[:, 2].astype(np.int8)
    assert s.o == len(s.b), "trailing bytes in symmetry dump"
    return out


def time_reference_parallel(work: str, wl, lam: np.ndarray, ksel, procs: int,
                            ekev: float | None = None) -> float:
    """Wall time of the reference beam loop over the directions ksel split into `procs`
    disjoint parts, one single-threaded driver process per part running at once, minus the
    time of one process on one direction (the per-energy setup every process repeats)."""
    parts = [p for p in np.array_split(np.asarray(ksel), procs) if len(p)]
    dirs = []
    for i in range(len(parts)):
        d = os.path.join(work, f"p{i}")
        os.makedirs(d, exist_ok=True)
        dirs.append(d)
    t0 = time.perf_counter()
    run_driver(dirs[0], wl, lam, [int(parts[0][0])], 2, ekev)
    t_setup = time.perf_counter() - t0
    from concurrent.futures import ThreadPoolExecutor
    t0 = time.perf_counter()
    with ThreadPoolExecutor(len(parts)) as ex:
        list(ex.map(lambda a: run_driver(a[0], wl, lam, list(a[1]), 2, ekev), zip(dirs, parts)))
    return (time.perf_counter() - t0) - t_setup


def have_reference() -> bool:
    return os.path.exists(os.path.join(REF_BIN, "ctemref_driver"))


def time_reference(work: str, wl, lam: np.ndarray, ksel, ekev: float | None = None) -> float:
    """Wall time of the reference beam loop (Compute_DynMat + CalcLgh3 per direction) over
    the directions ksel, single-threaded (MKL sequential), minus the time of the same run on
    one direction (which carries the per-energy setup and file output)."""
    t0 = time.perf_counter()
    run_driver(work, wl, lam, list(ksel), 2, ekev)
    t1 = time.perf_counter()
    run_driver(work, wl, lam, [int(ksel[0])], 2, ekev)
    t2 = time.perf_counter()
    return (t1 - t0) - (t2 - t1)


# ----------------------------------------------------------------------------------------
# The detector-pattern program (CTEMEBSD.f90), built as oracle/_ref/CTEMEBSD and, with the
# noise module's POIDEV replaced by the identity, oracle/_ref/CTEMEBSD_clean.

def write_mc_energy_file(path: str, accum_e, ehistmin: float, ebinsize: float, mcsig: float,
                         ekev: float = 20.0, xtalname: str = "crystal.xtal") -> None:
    """MC file (CTEMMC.f90:264-276) with energy histograms accum_e [2nsy+1][2nsx+1][numEbins]
    (C order == Fortran accum_e(numEbins, -nsx:nsx, -nsy:nsy)) and an empty depth array."""
    acc = np.ascontiguousarray(accum_e, np.int32)
    ny2, nx2, ne = acc.shape
    nzb = 2
    recs = [
        _rec(_chars("CTEMMC.f90", 132)),
        _rec(_chars("2.0/2013", 8)),
        _rec(_chars(xtalname, 132)),
        _rec(struct.pack("<4iqi", ne, nzb, nx2, ny2, int(acc.astype(np.int64).sum()), 1)),
        _rec(struct.pack("<5d", ekev, ehistmin, ebinsize, 100.0, 1.0)),
        _rec(struct.pack("<2d", mcsig, 0.0)),
        _rec(_chars("full", 4)),
        _rec(acc.tobytes()),
        _rec(np.zeros((ne, nzb, nx2 // 10 + 1, ny2 // 10 + 1), np.int32).tobytes()),
    ]
    with open(path, "wb") as f:
        f.write(b"".join(recs))


def run_ebsd_program(work: str, nml: dict, clean: bool, timeout: int = 300) -> np.ndarray:
    """Run CTEMEBSD (or CTEMEBSD_clean) in `work` on &EBSDdata `nml` (the angle, MC and master
    files already there); returns the output file's patterns float32 [n][ny][nx]."""
    def fmt(v):
        if isinstance(v, str):
            return f"'{v}'"
        if isinstance(v, (list, tuple)):
            return ", ".join(repr(float(x)) for x in v)
        return repr(v)
    with open(os.path.join(work, "CTEMEBSD.nml"), "w") as f:
        f.write(" &EBSDdata\n" + "".join(f" {k} = {fmt(v)},\n" for k, v in nml.items()) + " /\n")
    exe = os.path.join(REF_BIN, "CTEMEBSD_clean" if clean else "CTEMEBSD")
    subprocess.run(["timeout", "-k", "5", str(timeout), exe, "CTEMEBSD.nml"], cwd=work, check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    with open(os.path.join(work, nml.get("datafile", "EBSDout.data")), "rb") as f:
        data = f.read()
    recs, off = [], 0
    while off < len(data):
        (n,) = struct.unpack_from("<i", data, off)
        recs.append(data[off + 4: off + 4 + n])
        off += 8 + n
    nx, ny, num = struct.unpack("<3i", recs[0])
    return np.stack([np.frombuffer(r, np.float32).copy() for r in recs[1:1 + num]]), (nx, ny, num)


def run_ecp_program(work: str, xtal: dict, nml: dict, cutoffs=(4.0, 8.0, 0.05), timeout: int = 600) -> dict:
    """Run CTEMECP in `work` on &ECPlist `nml` (xtalname is set here) with the crystal `xtal` and
    the &BetheList cutoffs (weakcutoff, cutoff, sgcutoff); returns the output file's records
    (CTEMECP.f90:385-449) and the DynMat of the first direction the program writes to
    ECPmatrix.data (:301-306). The program's automatic arrays need an unlimited stack."""
    write_xtal(os.path.join(work, "crystal.xtal"), xtal)
    with open(os.path.join(work, "BetheParameters.nml"), "w") as f:
        f.write(" &BetheList\n weakcutoff = %r,\n cutoff = %r,\n sgcutoff = %r,\n /\n" % tuple(float(c) for c in cutoffs))

    def fmt(v):
        if isinstance(v, str):
            return f"'{v}'"
        if isinstance(v, (list, tuple)):
            return ", ".join(repr(x) for x in v)
        return repr(v)
    body = dict(nml, xtalname="crystal.xtal", outname="ecp.data")
    with open(os.path.join(work, "CTEMECP.nml"), "w") as f:
        f.write(" &ECPlist\n" + "".join(f" {k} = {fmt(v)},\n" for k, v in body.items()) + " /\n")
    exe = os.path.join(REF_BIN, "CTEMECP")
    subprocess.run(["bash", "-c", f"ulimit -s unlimited && exec timeout -k 5 {int(timeout)} {exe} CTEMECP.nml"],
                   cwd=work, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)

    def records(path):
        with open(path, "rb") as f:
            data = f.read()
        out, off = [], 0
        while off < len(data):
            (n,) = struct.unpack_from("<i", data, off)
            out.append(data[off + 4: off + 4 + n])
            off += 8 + n
        return out
    r = records(os.path.join(work, "ecp.data"))
    nx, ny, nt = struct.unpack("<3i", r[2])
    m = records(os.path.join(work, "ECPmatrix.data"))
    (nn,) = struct.unpack("<i", m[0])
    return dict(
        mode=struct.unpack("<i", r[4])[0], thetac=struct.unpack("<f", r[8])[0], ktmax=struct.unpack("<f", r[9])[0],
        numk=struct.unpack("<i", r[12])[0], ga=np.array(struct.unpack("<3i", r[14]), np.int32),
        galen=struct.unpack("<f", r[15])[0], delta=struct.unpack("<f", r[16])[0],
        gperp=np.frombuffer(r[17], np.float32).copy(), symlabels=np.frombuffer(r[18], np.int32).copy(),
        thick0=np.frombuffer(r[19], np.float32).copy(),
        sr=np.frombuffer(r[20], np.float32).reshape(nt, ny, nx).copy(),
        dynmat=np.frombuffer(m[1], np.complex128).reshape(nn, nn, order="F").copy())


def run_ecp_setup(work: str, xtal: dict, voltage: float, k, fn, dmin: float, thetac: float, npix: int,
                  cutoffs=(4.0, 8.0, 0.05), ndyn: int = 1, ktmax: float = 0.0, timeout: int = 300) -> dict:
    """Run oracle/_ref/ctemref_ecp (the channeling-pattern program's set-up, ref_ecp.f90) in `work`
    and parse its dump: the foil normal, ga/gb, ktmax/thetac, the reflection list with its LUT,
    the sites and orbits, every incident direction, and the beam sets and DynMat of the first
    `ndyn` directions."""
    write_xtal(os.path.join(work, "crystal.xtal"), xtal)
    with open(os.path.join(work, "BetheParameters.nml"), "w") as f:
        f.write(" &BetheList\n weakcutoff = %r,\n cutoff = %r,\n sgcutoff = %r,\n /\n" % tuple(float(c) for c in cutoffs))
    with open(os.path.join(work, "refecp.nml"), "w") as f:
        f.write(" &refecp\n xtalname = 'crystal.xtal',\n voltage = %r,\n k = %d, %d, %d,\n fn = %d, %d, %d,\n"
                " dmin = %r,\n thetac = %r,\n ktmax = %r,\n npix = %d,\n outname = 'ecpdump.bin',\n ndyn = %d,\n /\n"
                % (float(voltage), *[int(v) for v in k], *[int(v) for v in fn], float(dmin), float(thetac),
                   float(ktmax), int(npix), int(ndyn)))
    subprocess.run(["bash", "-c", f"ulimit -s unlimited && exec timeout -k 5 {int(timeout)} "
                    f"{os.path.join(REF_BIN, 'ctemref_ecp')}"], cwd=work, check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    with open(os.path.join(work, "ecpdump.bin"), "rb") as f:
        buf = f.read()
    off = 0

    def take(fmt):
        nonlocal off
        v = struct.unpack_from("<" + fmt, buf, off)
        off += struct.calcsize("<" + fmt)
        return v

    def arr(dtype, n):
        nonlocal off
        a = np.frombuffer(buf, dtype, n, off).copy()
        off += a.nbytes
        return a
    magic, numk, nref, ntype, sgnum = take("5i")
    assert magic == 20261
    ga = np.array(take("3i"), np.int32)
    gb = np.array(take("3i"), np.int32)
    dims = take("3i")
    lbs = take("3i")
    ktmax_, thetac_, bragg, galen = take("4f")
    dynfn = arr(np.float32, 3)
    mlambda, dynupz, vol = take("3d")
    dmt = arr(np.float64, 9).reshape(3, 3, order="F")
    rmt = arr(np.float64, 9).reshape(3, 3, order="F")
    ncell = dims[0] * dims[1] * dims[2]
    lut = arr(np.complex128, ncell)
    dbd = arr(np.int32, ncell)
    hkl = np.zeros((nref, 3), np.int32)
    ucg = np.zeros(nref, np.complex128)
    for i in range(nref):
        hkl[i] = take("3i")
        ucg[i] = complex(*take("2d"))
    sites = []
    for _ in range(ntype):
        z = take("i")[0]
        pos = arr(np.float32, 5)
        n = take("i")[0]
        orb = arr(np.float64, 3 * n).reshape(3, n).T.copy()
        sites.append(dict(Z=z, pos=pos, orbit=orb))
    ij = np.zeros((numk, 2), np.int32)
    kv = np.zeros((numk, 3))
    kt = np.zeros((numk, 3))
    kn = np.zeros(numk)
    for i in range(numk):
        ij[i] = take("2i")
        kv[i] = take("3d")
        kt[i] = take("3d")
        kn[i] = take("d")[0]
    dyn = []
    for _ in range(min(ndyn, numk)):
        nns, nnw = take("2i")
        shkl = arr(np.int32, 3 * nns).reshape(nns, 3)
        ssg = arr(np.float32, nns)
        D = arr(np.complex128, nns * nns).reshape(nns, nns, order="F")
        dyn.append(dict(nns=nns, nnw=nnw, hkl=shkl, sg=ssg, dynmat=D))
    assert off == len(buf)
    return dict(numk=numk, nref=nref, sgnum=sgnum, ga=ga, gb=gb, lut_dim=np.array(dims, np.int32),
                lut_lo=np.array(lbs, np.int32), ktmax=ktmax_, thetac=thetac_, bragg=bragg, galen=galen,
                dynfn=dynfn, mlambda=mlambda, dynupz=dynupz, vol=vol, dmt=dmt, rmt=rmt,
                lut=lut.reshape(dims, order="F"), dbdiff=dbd.reshape(dims, order="F"), hkl=hkl, ucg=ucg,
                sites=sites, ij=ij, k=kv, kt=kt, kn=kn, dyn=dyn)


def run_mc_program(work: str, xtal: dict, nml: dict, timeout: int = 600) -> dict:
    """Run CTEMMC (oracle/_ref/CTEMMC, src_programs/CTEMMC.f90 built from its sources) in `work` on
    &MCdata `nml` (xtalname and dataname are set here) and parse its output file
    (CTEMMC.f90:264-276): the header values and accum_e [numsy][numsx][numEbins],
    accum_z [ny][nx][numzbins][numEbins] (the Fortran arrays, energy fastest)."""
    write_xtal(os.path.join(work, "crystal.xtal"), xtal)

    def fmt(v):
        if isinstance(v, str):
            return f"'{v}'"
        return repr(v)
    body = dict(nml, xtalname="crystal.xtal", dataname="mc.data")
    with open(os.path.join(work, "CTEMMC.nml"), "w") as f:
        f.write(" &MCdata\n" + "".join(f" {k} = {fmt(v)},\n" for k, v in body.items()) + " /\n")
    subprocess.run(["timeout", "-k", "5", str(int(timeout)), os.path.join(REF_BIN, "CTEMMC"), "CTEMMC.nml"],
                   cwd=work, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return read_mc_file(os.path.join(work, "mc.data"))


def read_mc_file(path: str) -> dict:
    """The Monte Carlo output file of CTEMMC.f90:264-276."""
    with open(path, "rb") as f:
        data = f.read()
    recs, off = [], 0
    while off < len(data):
        (n,) = struct.unpack_from("<i", data, off)
        recs.append(data[off + 4: off + 4 + n])
        off += 8 + n
    numEbins, numzbins, numsx, numsy = struct.unpack_from("<4i", recs[3], 0)
    num_el, nthreads = struct.unpack_from("<qi", recs[3], 16)
    ekev, ehistmin, ebinsize, depthmax, depthstep = struct.unpack("<5d", recs[4])
    sig, omega = struct.unpack("<2d", recs[5])
    nx = (numsx - 1) // 2
    nz = 2 * (nx // 10) + 1
    acc_e = np.frombuffer(recs[7], np.int32).reshape(numsy, numsx, numEbins).copy()
    acc_z = np.frombuffer(recs[8], np.int32).reshape(nz, nz, numzbins, numEbins).copy()
    return dict(progname=recs[0].decode().strip(), version=recs[1].decode().strip(), xtalname=recs[2].decode().strip(),
                numEbins=numEbins, numzbins=numzbins, numsx=numsx, numsy=numsy, num_el=num_el, nthreads=nthreads,
                EkeV=ekev, Ehistmin=ehistmin, Ebinsize=ebinsize, depthmax=depthmax, depthstep=depthstep, sig=sig,
                omega=omega, MCmode=recs[6].decode(), accum_e=acc_e, accum_z=acc_z)


def run_mc_setup(work: str, xtal: dict, ekev: float, primeseed: int, ndraw: int = 1000, timeout: int = 120) -> dict:
    """oracle/_ref/ctemref_mc (ref_mc.f90): CalcDensity's density, avZ, avA (float32), the Lambert
    constants, the cell volume, numat and the first ndraw numbers of rng_uniform after
    rng_seed(primeseed + 1)."""
    write_xtal(os.path.join(work, "crystal.xtal"), xtal)
    with open(os.path.join(work, "refmc.nml"), "w") as f:
        f.write(" &refmc\n xtalname = 'crystal.xtal',\n EkeV = %r,\n primeseed = %d,\n ndraw = %d,\n"
                " outname = 'mcdump.bin',\n /\n" % (float(ekev), int(primeseed), int(ndraw)))
    subprocess.run(["timeout", "-k", "5", str(int(timeout)), os.path.join(REF_BIN, "ctemref_mc")], cwd=work,
                   check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    with open(os.path.join(work, "mcdump.bin"), "rb") as f:
        buf = f.read()
    magic, nd, ntype = struct.unpack_from("<3i", buf, 0)
    assert magic == 20262
    dens, avz, ava = np.frombuffer(buf, np.float32, 3, 12)
    spio2, spi2, vol = struct.unpack_from("<3d", buf, 24)
    numat = np.frombuffer(buf, np.int32, ntype, 48).copy()
    u = np.frombuffer(buf, np.float32, nd, 48 + 4 * ntype).copy()
    return dict(dens=dens, avZ=avz, avA=ava, sPio2=spio2, sPi2=spi2, vol=vol, numat=numat, u=u)
