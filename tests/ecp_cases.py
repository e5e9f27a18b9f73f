"""Shared helpers of the channeling-pattern tests: the set-ups and runs of tests/golden/ecp.npz (the
reference program src_programs/CTEMECP.f90 and its set-up, dumped by oracle/ref_ecp.f90)."""
import ast

import numpy as np

from conftest import load_golden


def fixture():
    return load_golden("ecp")


def run_params(z, run):
    return ast.literal_eval(bytes(z[f"{run}_run"]).decode())


def setup_params(z, key):
    volt, dmin, th, npix, wc, c, sgc = (float(v) for v in z[f"{key}_params"])
    return dict(xtal=ast.literal_eval(bytes(z[f"{key}_xtal"]).decode()), voltage=volt, dmin=dmin, thetac=th,
                npix=int(npix), cutoffs=(wc, c, sgc), k=[int(v) for v in z[f"{key}_zone"]],
                fn=[int(v) for v in z[f"{key}_fn"]])


def thicknesses(rp):
    return (np.float32(rp["startthick"]) + np.float32(rp["thickinc"]) *
            np.arange(rp["numthick"], dtype=np.float32)).astype(np.float32)


def oracle_crystal(z, key):
    """oracle.EcpCrystal from the reference's set-up dump"""
    from oracle import oracle as O
    sp = setup_params(z, key)
    lo = np.asarray(z[f"{key}_lut_lo"])
    hkl = np.asarray(z[f"{key}_hkl"])
    h, k, l = (hkl - lo).T
    dbd = np.asarray(z[f"{key}_dbdiff"])[h, k, l]
    occ = [s["occ"] for s in sp["xtal"]["sites"]]
    sites = [dict(Z=int(z[f"{key}_site{i}_Z"]), B=float(z[f"{key}_site{i}_pos"][4]), occ=float(occ[i]),
                  orbit=np.asarray(z[f"{key}_site{i}_orbit"])) for i in range(int(z[f"{key}_nsite"]))]
    mlambda, dynupz, _ = (float(v) for v in z[f"{key}_scalars"])
    lut = np.asarray(z[f"{key}_lut"]).astype(np.complex128).transpose(2, 1, 0)
    return O.EcpCrystal(hkl, np.asarray(z[f"{key}_ucg"]).astype(np.complex128), dbd, lut, lo, mlambda, dynupz,
                        sp["cutoffs"], z[f"{key}_rmt"], sites)


def native_setup(z, key):
    """The native set-up of the same crystal and parameters (ctem_setup_create_ecp and the
    geometry entries of ctem_ecp.h): (ebsd.Setup, values, dict of the geometry)"""
    from ctemsoft_amd import ecp, xtal
    sp = setup_params(z, key)
    st0, v0 = xtal.native_setup(sp["xtal"], sp["voltage"] / 1000.0, sp["dmin"], sp["cutoffs"],
                                ecp=(sp["voltage"], sp["k"], 0.0))
    dynfn = ecp.foil_normal(sp["fn"], v0["dmt"], v0["rmt"])
    ga = ecp.shortest_g(sp["k"], v0["dmt"], v0["rmt"])
    thetac, ktmax, bragg, galen = ecp.convergence(ga, v0["rmt"], v0["mLambda"], sp["thetac"], 0.0)
    st, vals = xtal.native_setup(sp["xtal"], sp["voltage"] / 1000.0, sp["dmin"], sp["cutoffs"],
                                 ecp=(sp["voltage"], sp["k"], np.float32(thetac) / np.float32(1000.0)))
    kv = ecp.kvectors(sp["k"], ga, ktmax, sp["npix"], v0["dmt"], v0["rmt"], v0["mLambda"])
    geo = dict(dynfn=dynfn, ga=ga, conv=np.array([ktmax, thetac, bragg, galen], np.float32), **kv)
    return st, vals, geo
