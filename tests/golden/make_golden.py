"""Generate the golden fixtures under tests/golden/ from the compiled reference.

Runs in the development container only (it needs /root/reference and the binaries that
`make -C oracle ref` builds into oracle/_ref/). It writes the crystal (.xtal), Monte Carlo
depth file and namelists for each workload, runs

  * oracle/_ref/ctemref_driver  -- the reference's own modules plus its CalcLgh3, dumping
    LUT, reflection list, beam sets, DynMat and Lgh for selected incident directions;
  * oracle/_ref/CTEMEBSDmaster  -- the reference program itself, end to end, for the
    small fcc-Ni workload (its float32 master pattern is kept whole),

and stores the numbers as compressed .npz fixtures (data only).

File layouts written here follow the reference readers:
  .xtal       files.f90:271-277 (two Fortran sequential records)
  MC file     CTEMMC.f90:264-276 / CTEMEBSDmaster.f90:190-235
  master file CTEMEBSDmaster.f90:540-568
"""
from __future__ import annotations

import argparse
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, REPO)

from ctemsoft_amd import workloads  # noqa: E402
from oracle.refrun import read_master, run_driver, run_program, run_sym  # noqa: E402

def _setup_arrays(d: dict) -> dict:
    out = dict(isym=d["isym"], usehex=d["usehex"], numk=d["numk"], sgnum=d["sgnum"],
               mLambda=d["mLambda"], mRelcor=d["mRelcor"], mPsihat=d["mPsihat"],
               mSigma=d["mSigma"], DynUpz=d["DynUpz"], cutoffs=np.array(
                   [d["weakcutoff"], d["cutoff"], d["sgcutoff"]]), vol=d["vol"],
               dmt=d["dmt"], rmt=d["rmt"], dsm=d["dsm"], rsm=d["rsm"],
               lut=d["lut"].astype(np.complex64), dbdiff=d["dbdiff"].astype(np.uint8),
               lut_lo=d["lut_dims"][3:6], hkl=d["hkl"].astype(np.int16),
               ucg=d["ucg"].astype(np.complex64), kij=d["kij"], kvec=d["kvec"], kn=d["kn"],
               thick=d["thick"])
    assert np.array_equal(out["lut"].astype(np.complex128), d["lut"]), "LUT holds non-f32 values"
    for i, st in enumerate(d["sites"]):
        out[f"site{i}_Z"] = st["Z"]
        out[f"site{i}_pos"] = st["pos"]
        out[f"site{i}_orbit"] = st["orbit"]
    out["nsite"] = len(d["sites"])
    return out


def _beams_csr(per) -> dict:
    ns = np.array([len(p["strong"]) for p in per], np.int32)
    nw = np.array([len(p["weak"]) for p in per], np.int32)
    return dict(ik=np.array([p["ik"] for p in per], np.int32), nns=ns, nnw=nw,
                strong=np.concatenate([p["strong"] for p in per]).astype(np.int32),
                strongsg=np.concatenate([p["strongsg"] for p in per]).astype(np.float32),
                weak=np.concatenate([p["weak"] for p in per] + [np.zeros(0, np.int32)]).astype(np.int32),
                weaksg=np.concatenate([p["weaksg"] for p in per] + [np.zeros(0, np.float32)]).astype(np.float32))


def golden_workload(name: str, nsel_dyn: int, nsel_beams: int | None, full_program: bool,
                    ekev: float | None = None, tag: str | None = None,
                    corners: bool = True, lutfull: bool = False) -> str:
    wl = workloads.get(name)
    lam, counts, tail = wl.depth_weights_for_program(ekev)
    with tempfile.TemporaryDirectory(prefix="ctemgold_") as work:
        setup = run_driver(work, wl, lam, [0], 0, ekev, lutfull=lutfull)
        numk = setup["numk"]
        rng = np.random.default_rng(wl.seed)
        # beam sets: all directions, or a seeded subset
        if nsel_beams is None or nsel_beams >= numk:
            kb = np.arange(numk)
        else:
            kb = np.sort(rng.choice(numk, nsel_beams, replace=False))
        beams = run_driver(work, wl, lam, kb, 0, ekev)
        # DynMat + Lgh on a small seeded subset that always holds the zone axis and (unless
        # the matrices are large) the last direction of the list
        special = {0, numk - 1} if corners else {0}
        kd = np.sort(np.array(sorted(special | set(rng.choice(numk, nsel_dyn, replace=False).tolist()))))
        dyn = run_driver(work, wl, lam, kd, 2, ekev)
        out = _setup_arrays(setup)
        if lutfull:
            # CalcUcg over the box that holds every difference of two list entries (lattice-
            # allowed g only). Inside the reference's own LUT box the two agree except where
            # the reference also stored family members that the centring rule forbids
            # (|Ucg| at round-off level); the reference's values are kept there.
            lf = setup["lut_full"].copy()
            lo, flo = setup["lut_dims"][3:6], setup["lut_full_lo"]
            d = setup["lut"].shape
            o = lo - flo
            inner = lf[o[2]:o[2] + d[0], o[1]:o[1] + d[1], o[0]:o[0] + d[2]]
            diff = inner != setup["lut"]
            assert np.all(inner[diff] == 0)
            assert np.abs(setup["lut"][diff]).max() < 1e-6 * np.abs(setup["lut"]).max()
            inner[...] = setup["lut"]
            out["lut_full"] = lf.astype(np.complex64)
            assert np.array_equal(out["lut_full"].astype(np.complex128), lf)
            out["lut_full_lo"] = flo
        out["lam"] = lam
        out["dz"] = wl.dz_for(ekev)
        out["ekev"] = wl.ekev if ekev is None else ekev
        for k, v in _beams_csr(beams["per"]).items():
            out["beams_" + k] = v
        out["dyn_ik"] = np.array([p["ik"] for p in dyn["per"]], np.int32)
        for j, p in enumerate(dyn["per"]):
            out[f"dyn{j}_dynmat"] = p["dynmat"]
            out[f"dyn{j}_lgh"] = p["lgh"]
        if full_program:
            m = run_program(work, wl, counts, tail)
            out["master_sr"] = m["sr"]
            out["master_grid"] = m["grid"]
    fn = os.path.join(HERE, f"{tag or name}.npz")
    np.savez_compressed(fn, **out)
    return fn


def setup_series(name: str) -> str:
    """Per-energy setup (LUT, reflection list, sites, wavelength) of every energy bin of an
    energy-series workload, from the reference's own modules; one archive, keys e<i>_<field>."""
    wl = workloads.get(name)
    out = {}
    with tempfile.TemporaryDirectory(prefix="ctemser_") as work:
        for i, e in enumerate(wl.energies):
            lam, _, _ = wl.depth_weights_for_program(e)
            d = run_driver(work, wl, lam, [0], 0, e)
            for k, v in _setup_arrays(d).items():
                if k in ("kij", "kvec", "kn"):
                    continue  # the library computes the directions
                out[f"e{i}_{k}"] = v
            out[f"e{i}_ekev"] = e
            out[f"e{i}_lam"] = lam
            out[f"e{i}_dz"] = wl.dz_for(e)
    out["nenergy"] = len(wl.energies)
    fn = os.path.join(HERE, f"{name}_setups.npz")
    np.savez_compressed(fn, **out)
    return fn


def golden_sym() -> str:
    """Laue-equivalence table (isym 1..13, pixels -16..16 squared) and the RoscaLambert
    direction lists of every Laue group on the square and hexagonal grids (npx 11), from the
    reference's Apply2DLaueSymmetry and Calckvectors. The structure matrix is corundum's
    (hexagonal axes) and the wavelength that of 15 kV; both enter only AddkVector's final
    transformation."""
    fx = np.load(os.path.join(HERE, "al2o3_c2.npz"))
    with tempfile.TemporaryDirectory(prefix="ctemsym_") as work:
        out = run_sym(work, 16, 11, fx["dsm"], float(fx["mLambda"]))
    out["dsm"] = fx["dsm"]
    out["mLambda"] = fx["mLambda"]
    out["nbox"] = 16
    out["npx"] = 11
    fn = os.path.join(HERE, "sym_tables.npz")
    np.savez_compressed(fn, **out)
    return fn


def golden_ebsd() -> str:
    """Detector patterns from the reference's CTEMEBSD program (its output file) and from the same
    program with POIDEV replaced by the identity (the noise-free patterns), for two cases on a
    64 x 48 detector: (a) Euler angles, raw output; (b) quaternions with the optional axis-angle
    rotation, 2 x 2 binning and gamma, a three-site master (the program's site sum). The master is
    the reference program's fcc-Ni master (ni_c0) extended to three energy bins with seeded
    perturbations; the MC histograms are a seeded smooth field. The energy window covers every
    bin: with a narrower window the program sums unassigned memory into num_el
    (CTEMEBSD.f90:470-508) and its output is undefined."""
    from oracle.refrun import run_ebsd_program, write_mc_energy_file
    from ctemsoft_amd.masterfile import write_master
    ni = np.load(os.path.join(HERE, "ni_c0.npz"))
    sr0 = np.asarray(ni["master_sr"], np.float32)  # [1][1][2npy+1][2npx+1]
    rng = np.random.default_rng(20260312)
    nE, nsx = 3, 20
    out = {}
    for case in ("a", "b"):
        nset = 1 if case == "a" else 3
        sr = np.concatenate([sr0 * np.float32(0.8 + 0.2 * e) for e in range(nE)], axis=1)
        sr = np.concatenate([sr * np.float32(1.0 + 0.37 * q) for q in range(nset)], axis=0)
        sr = (sr + rng.random(sr.shape, dtype=np.float32) * np.float32(0.01)).astype(np.float32)
        xs = np.arange(-nsx, nsx + 1)
        X, Y = np.meshgrid(xs, xs, indexing="xy")  # [y][x]
        base = np.exp(-((X - 3) ** 2 + (Y + 2) ** 2) / 300.0)
        acc = np.stack([np.round(base * (4000 * (e + 1)) + rng.integers(0, 50, base.shape)) for e in range(nE)],
                       axis=-1).astype(np.int32)  # [y][x][E]
        if case == "a":
            ang = np.array([[10.0, 20.0, 30.0], [120.5, 45.25, 300.0], [0.0, 0.0, 0.0], [250.0, 80.0, 33.3]], np.float32)
            nml = dict(L=15000.0, thetac=10.0, delta=50.0, numsx=64, numsy=48, xpc=3.5, ypc=-2.25,
                       anglefile="angles.txt", eulerconvention="tsl", FZfile="master.data", energyfile="mc.data",
                       datafile="ebsdout.data", beamcurrent=14.513, dwelltime=100.0, energymin=17.6,
                       energymax=20.4, binning=1, gammavalue=1.0, scalingmode="not", axisangle=[0.0, 0.0, 0.0, 0.0])
            kind = "eu"
        else:
            q = rng.normal(size=(5, 4)).astype(np.float64)
            q[:, 0] = np.abs(q[:, 0])
            ang = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
            nml = dict(L=18000.0, thetac=6.5, delta=40.0, numsx=64, numsy=48, xpc=-1.25, ypc=2.5,
                       anglefile="angles.txt", eulerconvention="tsl", FZfile="master.data", energyfile="mc.data",
                       datafile="ebsdout.data", beamcurrent=20.0, dwelltime=50.0, energymin=18.0,
                       energymax=20.0, binning=2, gammavalue=0.7, scalingmode="gam",
                       axisangle=[0.0, 0.6, 0.8, 12.5])
            kind = "qu"
        res = {}
        with tempfile.TemporaryDirectory(prefix="ctemdet_") as work:
            write_mc_energy_file(os.path.join(work, "mc.data"), acc, 18.0, 1.0, 70.0)
            write_master(os.path.join(work, "master.data"), sr, [18.0, 19.0, 20.0], [28] * nset)
            with open(os.path.join(work, "angles.txt"), "w") as f:
                f.write(f"{kind}\n{len(ang)}\n")
                for r in ang:
                    f.write(" ".join(repr(float(v)) for v in r) + "\n")
            res["noisy"], hdr = run_ebsd_program(work, nml, clean=False)
            res["clean"], _ = run_ebsd_program(work, nml, clean=True)
            with open(os.path.join(work, "angles.txt")) as f:
                angles_txt = f.read()
        out[f"{case}_sr"] = sr
        out[f"{case}_accum_e"] = acc
        out[f"{case}_angles"] = ang
        out[f"{case}_angles_txt"] = np.frombuffer(angles_txt.encode(), np.uint8)
        out[f"{case}_kind"] = kind
        out[f"{case}_nml"] = np.frombuffer(repr(nml).encode(), np.uint8)
        out[f"{case}_header"] = np.array(hdr, np.int32)
        out[f"{case}_noisy"] = res["noisy"]
        out[f"{case}_clean"] = res["clean"]
    out["Ehistmin"], out["Ebinsize"], out["MCsig"], out["ekevs"] = 18.0, 1.0, 70.0, np.array([18.0, 19.0, 20.0], np.float32)
    fn = os.path.join(HERE, "ebsd_det.npz")
    np.savez_compressed(fn, **out)
    return fn


GAAS = dict(system=1, abc=(0.56533, 0.56533, 0.56533), angles=(90.0, 90.0, 90.0), sg=216, sgset=1,
            sites=[dict(Z=31, xyz=(0.0, 0.0, 0.0), occ=1.0, B=0.0060), dict(Z=33, xyz=(0.25, 0.25, 0.25), occ=0.95, B=0.0050)])

# channeling-pattern set-ups (crystal, voltage [V], zone axis, foil normal, dmin, thetac, npix,
# Bethe cutoffs) and runs (set-up, compmode, startthick, thickinc, numthick, zintstep)
ECP_SETUPS = {
    "s1": ("ni", 20000.0, (0, 0, 1), (0, 0, 1), 0.05, 120.0, 6, (4.0, 8.0, 0.05)),
    "s2": ("gaas", 15000.0, (0, 1, 1), (0, 1, 2), 0.06, 90.0, 5, (20.0, 40.0, 0.05)),
    "s3": ("ni", 20000.0, (1, 1, 2), (1, 1, 3), 0.05, 100.0, 3, (4.0, 8.0, 0.05)),
}
ECP_RUNS = {
    "a": ("s1", "Blochwv", 2.0, 2.0, 10, 1.0),
    "b": ("s1", "ScatMat", 2.0, 2.0, 10, 1.0),
    "c": ("s2", "ScatMat", 3.0, 2.5, 4, 0.5),
    "d": ("s2", "Blochwv", 3.0, 2.0, 6, 0.5),
    "e": ("s3", "ScatMat", 0.1, 1.0, 5, 0.25),   # nint(0.1/0.25) = 0: no thickness is ever stored
    "f": ("s3", "Blochwv", 0.5, 7.5, 3, 1.0),
}


def golden_ecp() -> str:
    """The channeling-pattern program (src_programs/CTEMECP.f90, oracle/_ref/CTEMECP) on three
    set-ups -- fcc Ni [001], zinc-blende GaAs [011] with foil normal [012] and 95% As occupancy,
    Ni [112] with foil normal [113] -- in both solution modes: the output file's records (sr on
    the disk's directions, ga, ktmax, galen, delta, gperp, the symmetry labels) and the first
    direction's DynMat (ECPmatrix.data); and the set-up driver (oracle/ref_ecp.f90) dump: foil
    normal, ga, convergence, reflection list with its LUT, sites, and every incident direction."""
    from oracle.refrun import run_ecp_program, run_ecp_setup
    from ctemsoft_amd import workloads
    xtals = dict(ni=workloads.get("ni_c0").xtal, gaas=GAAS)
    out = {}
    setups = {}
    with tempfile.TemporaryDirectory(prefix="ctemecp_") as work:
        for key, (xn, volt, k, fn, dmin, th, npix, cut) in ECP_SETUPS.items():
            d = run_ecp_setup(work, xtals[xn], volt, k, fn, dmin, th, npix, cutoffs=cut, ndyn=1)
            setups[key] = d
            out[f"{key}_xtal"] = np.frombuffer(repr(xtals[xn]).encode(), np.uint8)
            out[f"{key}_params"] = np.array([volt, dmin, th, npix, *cut], np.float64)
            out[f"{key}_zone"] = np.array(k, np.int32)
            out[f"{key}_fn"] = np.array(fn, np.int32)
            for f in ("ga", "dynfn", "hkl", "lut_lo", "ij", "kt", "kn", "dmt", "rmt"):
                out[f"{key}_{f}"] = np.asarray(d[f])
            out[f"{key}_kvec"] = np.asarray(d["k"])
            out[f"{key}_conv"] = np.array([d["ktmax"], d["thetac"], d["bragg"], d["galen"]], np.float32)
            out[f"{key}_scalars"] = np.array([d["mlambda"], d["dynupz"], d["vol"]], np.float64)
            out[f"{key}_ucg"] = d["ucg"].astype(np.complex64)
            out[f"{key}_lut"] = d["lut"].astype(np.complex64)          # Fortran LUT(h,k,l), float32 values
            out[f"{key}_dbdiff"] = d["dbdiff"].astype(np.uint8)
            for i, st in enumerate(d["sites"]):
                out[f"{key}_site{i}_Z"] = np.int32(st["Z"])
                out[f"{key}_site{i}_pos"] = st["pos"]
                out[f"{key}_site{i}_orbit"] = st["orbit"]
            out[f"{key}_nsite"] = np.int32(len(d["sites"]))
            dy = d["dyn"][0]
            out[f"{key}_dyn0_hkl"] = dy["hkl"]
            out[f"{key}_dyn0_sg"] = dy["sg"]
            out[f"{key}_dyn0_nnw"] = np.int32(dy["nnw"])
            out[f"{key}_dyn0_dynmat"] = dy["dynmat"]
        for run, (key, mode, t0, dt, nt, zint) in ECP_RUNS.items():
            xn, volt, k, fn, dmin, th, npix, cut = ECP_SETUPS[key]
            nml = dict(voltage=volt, k=list(k), fn=list(fn), dmin=dmin, thetac=th, startthick=t0, thickinc=dt,
                       numthick=nt, npix=npix, compmode=mode, zintstep=zint)
            r = run_ecp_program(work, xtals[xn], nml, cutoffs=cut)
            d = setups[key]
            assert r["numk"] == d["numk"] and np.array_equal(r["ga"], d["ga"])
            assert np.array_equal(r["dynmat"], d["dyn"][0]["dynmat"])
            ij = d["ij"]
            # sr on the disk's directions only: the program leaves the pixels outside unassigned
            out[f"{run}_sr"] = np.stack([r["sr"][:, j + npix, i + npix] for i, j in ij], axis=1)
            out[f"{run}_run"] = np.frombuffer(repr(dict(setup=key, compmode=mode, startthick=t0, thickinc=dt,
                                                         numthick=nt, zintstep=zint)).encode(), np.uint8)
            out[f"{run}_header"] = np.array([r["mode"], r["numk"]], np.int32)
            out[f"{run}_file"] = np.array([r["thetac"], r["ktmax"], r["galen"], r["delta"], *r["gperp"]], np.float32)
            out[f"{run}_symlabels"] = r["symlabels"]
    fn = os.path.join(HERE, "ecp.npz")
    np.savez_compressed(fn, **out)
    return fn


# Monte Carlo runs: the program's arrays must come from fresh (zeroed) memory -- it never
# initialises them -- so every histogram here is above the C library's 128 KiB mmap threshold;
# numsx keeps (numsx - 1) / 2 mod 10 < 5, where the program's depth map stays inside its array
MC_RUNS = {
    "m1": ("ni", dict(sig=70.0, omega=0.0, numsx=101, primeseed=932117, num_el=20000, nthreads=1, EkeV=20.0,
                      Ehistmin=10.0, Ebinsize=1.0, depthmax=100.0, depthstep=1.0)),
    "m2": ("fe", dict(sig=65.0, omega=7.5, numsx=81, primeseed=4711, num_el=10000, nthreads=1, EkeV=15.0,
                      Ehistmin=5.0, Ebinsize=0.5, depthmax=60.0, depthstep=2.0)),
    "m3": ("gaas", dict(sig=75.0, omega=-5.0, numsx=121, primeseed=123457, num_el=8000, nthreads=1, EkeV=25.0,
                        Ehistmin=15.0, Ebinsize=1.0, depthmax=50.0, depthstep=1.0)),
}


def golden_mc() -> str:
    """The Monte Carlo program (src_programs/CTEMMC.f90, oracle/_ref/CTEMMC) on three runs -- fcc Ni,
    bcc Fe with the second tilt omega, zinc-blende GaAs with 95% As occupancy -- its whole output
    file; and the set-up driver (oracle/ref_mc.f90): CalcDensity's values, numat and the first 2000
    numbers of the program's random stream for each run's seed."""
    from oracle.refrun import run_mc_program, run_mc_setup
    from ctemsoft_amd import workloads
    xtals = dict(ni=workloads.get("ni_c0").xtal, fe=workloads.get("fe_c1").xtal, gaas=GAAS)
    out = {}
    with tempfile.TemporaryDirectory(prefix="ctemmc_") as work:
        for run, (xn, nml) in MC_RUNS.items():
            d = run_mc_setup(work, xtals[xn], nml["EkeV"], nml["primeseed"], 2000)
            r = run_mc_program(work, xtals[xn], nml)
            with open(os.path.join(work, "mc.data"), "rb") as f:
                out[f"{run}_file"] = np.frombuffer(f.read(), np.uint8)
            out[f"{run}_xtal"] = np.frombuffer(repr(xtals[xn]).encode(), np.uint8)
            out[f"{run}_nml"] = np.frombuffer(repr(nml).encode(), np.uint8)
            out[f"{run}_density"] = np.array([d["dens"], d["avZ"], d["avA"]], np.float32)
            out[f"{run}_numat"] = d["numat"]
            out[f"{run}_u"] = d["u"]
            out[f"{run}_lambert"] = np.array([d["sPio2"], d["sPi2"], d["vol"]])
    fn = os.path.join(HERE, "mc.npz")
    np.savez_compressed(fn, **out)
    return fn


def golden_south() -> str:
    """The southern hemisphere of the corundum workload with the wide cutoffs (al2o3_c2w). The
    reference program computes the northern hemisphere only; the library's southern half takes
    the same Lambert points with the Cartesian z negated before the crystal-frame transform
    (ctem_kvectors, hemisphere -1). Here the reference's own beam loop body (ref_driver.f90, given
    those wave vectors) on 32 seeded southern directions and on the two (list indices 23995 and
    23996, 0-based) where a weak beam lies within ~3e-6 of the Bragg condition, so that the Bethe
    terms blow up: its beam sets on all, and on the two its DynMat maximum and whether CalcLgh3's
    Lgh is finite, also for their northern twins."""
    from ctemsoft_amd import ebsd
    wl = workloads.get("al2o3_c2w")
    lam, _, _ = wl.depth_weights_for_program()
    blow = np.array([23995, 23996])
    out = {}
    with tempfile.TemporaryDirectory(prefix="ctemsouth_") as work:
        setup = run_driver(work, wl, lam, [0], 0)
        numk = setup["numk"]
        npx, npy = ebsd.lambert_dims(wl.npx, setup["usehex"])
        ij, k, kn = ebsd.kvectors(npx, npy, setup["isym"], setup["usehex"], setup["dsm"], setup["mLambda"], -1)
        assert len(k) == numk
        rng = np.random.default_rng(17)
        kb = np.sort(np.array(sorted(set(rng.choice(numk, 32, replace=False).tolist()) | set(blow.tolist()))))
        beams = run_driver(work, wl, lam, None, 0, kvec=(k[kb], kn[kb]))
        per = beams["per"]
        for j, p in enumerate(per):
            p["ik"] = int(kb[j])
        for key, v in _beams_csr(per).items():
            out["beams_" + key] = v
        out["kvec"], out["kn"], out["kij"] = k[kb], kn[kb], ij[kb]
        for tag, kk in (("south", (k[blow], kn[blow])), ("north", (setup["kvec"][blow], setup["kn"][blow]))):
            d = run_driver(work, wl, lam, None, 2, kvec=kk)
            out[f"blow_{tag}_lgh_finite"] = np.array([np.isfinite(p["lgh"]).all() for p in d["per"]])
            out[f"blow_{tag}_dynmat_max"] = np.array([np.abs(p["dynmat"]).max() for p in d["per"]])
            out[f"blow_{tag}_min_sw"] = np.array([np.abs(p["weaksg"]).min() for p in d["per"]], np.float32)
        out["blow_ik"] = blow.astype(np.int32)
        out["numk"] = np.int32(numk)
    fn = os.path.join(HERE, "al2o3_c2w_south.npz")
    np.savez_compressed(fn, **out)
    return fn


def golden_refdots() -> str:
    """The reference's intensities on many directions, for the FP64 path's parity beyond the few
    directions whose Lgh the workload fixtures hold: the reference's own DynMat and CalcLgh3
    (ref_driver.f90, eigen-based) on 128 seeded directions of fe_c1w and 64 of fe_c3w at 10 kV,
    each Lgh contracted here with the oracle's S_p (its per-site structure matrix, pinned to the
    reference through the fcc-Ni master pattern) to dots_p = Re sum_gh Lgh(g,h) S_p(g,h); only the
    dots are stored."""
    from oracle import oracle as O
    out = {}
    for name, fixture, count, ekev in (("fe_c1w", "fe_c1w", 128, None), ("fe_c3w", "fe_c3w_10.0kV", 64, 10.0)):
        wl = workloads.get(name)
        fx = np.load(os.path.join(HERE, f"{fixture}.npz"))
        cr = O.Crystal(fx)
        lam, _, _ = wl.depth_weights_for_program(ekev)
        assert np.array_equal(lam, fx["lam"])
        rng = np.random.default_rng(23)
        ks = np.sort(rng.choice(len(fx["kvec"]), count, replace=False))
        with tempfile.TemporaryDirectory(prefix="ctemrefdots_") as work:
            d = run_driver(work, wl, lam, ks, 2, ekev)
        nsite = int(fx["nsite"])
        dots = np.zeros((len(ks), nsite))
        for j, p in enumerate(d["per"]):
            assert p["ik"] == ks[j]
            for site in range(nsite):
                dots[j, site] = float(np.real(np.sum(p["lgh"] * O.sgh(cr, p["strong"], site))))
        out[f"{fixture}_ik"] = ks.astype(np.int32)
        out[f"{fixture}_dots"] = dots
    fn = os.path.join(HERE, "ref_dots.npz")
    np.savez_compressed(fn, **out)
    return fn


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", nargs="*", default=["ni_c0", "ni_c0w", "fe_c1", "fe_c1w"])
    a = ap.parse_args()
    for w in a.which:
        if w in ("ni_c0", "ni_c0w"):
            print(golden_workload(w, nsel_dyn=5, nsel_beams=None, full_program=True))
        elif w in ("fe_c1", "fe_c1w"):
            print(golden_workload(w, nsel_dyn=3, nsel_beams=256, full_program=False))
        elif w == "al2o3_c2":
            print(golden_workload(w, nsel_dyn=5, nsel_beams=256, full_program=False, lutfull=True))
        elif w == "al2o3_c2w":
            print(golden_workload(w, nsel_dyn=1, nsel_beams=64, full_program=False,
                                  corners=False))
        elif w in ("fe_c3", "fe_c3w"):
            for e in (10.0, 15.0, 20.0):
                print(golden_workload(w, nsel_dyn=0, nsel_beams=64, full_program=False,
                                      ekev=e, tag=f"{w}_{e:.1f}kV", corners=False))
            if w == "fe_c3w":
                print(setup_series(w))
        elif w == "al2o3_c2x":
            # matrices past 512 beams: beam sets on 32 directions, the reference's Lgh (zgeev) on
            # the two largest of 3 seeded directions, kept as upper triangles (Lgh is Hermitian);
            # the LUT and DynMat are al2o3_c2's and the oracle's (both pinned to the reference)
            fn = golden_workload(w, nsel_dyn=2, nsel_beams=32, full_program=False, corners=False)
            d = dict(np.load(fn))
            sizes = [d[f"dyn{j}_lgh"].shape[0] for j in range(len(d["dyn_ik"]))]
            keep = [j for j in np.argsort(sizes)[::-1][:2]]
            out = {k: v for k, v in d.items() if not k.startswith("dyn") and k not in ("lut", "dbdiff")}
            out["dyn_ik"] = d["dyn_ik"][keep]
            for jj, j in enumerate(keep):
                L = d[f"dyn{j}_lgh"]
                iu = np.triu_indices(L.shape[0])
                out[f"dyn{jj}_lgh_n"] = np.int32(L.shape[0])
                out[f"dyn{jj}_lgh_upper"] = L[iu]
            np.savez_compressed(fn, **out)
            print(fn, [d[f"dyn{j}_lgh"].shape[0] for j in keep])
        elif w == "sym":
            print(golden_sym())
        elif w == "ebsd":
            print(golden_ebsd())
        elif w == "ecp":
            print(golden_ecp())
        elif w == "mc":
            print(golden_mc())
        elif w == "south":
            print(golden_south())
        elif w == "refdots":
            print(golden_refdots())
        else:
            raise SystemExit(f"unknown workload {w}")


if __name__ == "__main__":
    main()
