"""Complex-step derivatives of one direction's dots = Re sum_gh Lgh(g,h) S_p(g-h) with respect to
real parameters -- the real and imaginary part of single look-up-table entries U(d), and the site
Debye-Waller factors B_p. TEST INFRASTRUCTURE ONLY: the reference for the library's reverse mode
(ctem_plan_adjoint); never used by the product.

Every complex quantity of the forward path is carried as a pair of real arrays (re, im), and the
matrix exponential acts on the real 2n x 2n form [[Re A, -Im A], [Im A, Re A]], so that each
operation is an analytic function of the real parameters. A parameter x then takes the step
x + i h (h = 1e-30) in complex arithmetic, and

    d dots / dx = Im dots(x + i h) / h

holds to machine precision with no cancellation (Squire and Trapp, SIAM Review 40 (1998) 110).

The forward restates the oracle (oracle/ctem_oracle.c, itself pinned to the reference):
  DynMat, Bethe-corrected (gvectors.f90:1123-1173), with the beam sets held fixed;
  A = 2 pi i dz DynMat / fl32(2 kn), E = expm(A), psi_k = E^k e_1, Lgh = (dz/thick) sum_k
  lam_k psi_k psi_k^H (CTEMEBSDmaster.f90:610-722);
  S_p(d) = sum_orbit Z^2 exp(-B |d|^2/4) e^{2 pi i d.r} (:418-449) as the smooth function whose
  float32 terms the reference rounds (the derivative of the rounded terms is zero almost
  everywhere; the reverse mode differentiates the smooth function, as this does).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg

H = 1e-30


def _lut_lookup(lut, lo, dims, d):
    """LUT entries at integer offsets d[..., 3] (Fortran order array lut[l][k][h]); returns the
    flat cell index and a validity mask."""
    h = d[..., 0] - lo[0]
    k = d[..., 1] - lo[1]
    l = d[..., 2] - lo[2]
    cell = h + dims[0] * (k + dims[1] * l)
    ok = (h >= 0) & (h < dims[0]) & (k >= 0) & (k < dims[1]) & (l >= 0) & (l < dims[2])
    return cell, ok


def dots_forward(setup: dict, beams, kn, lam, dz, thick, lut_re, lut_im, site_b):
    """dots[p] of one direction for LUT real/imaginary parts lut_re/lut_im (flat, complex dtype
    allowed) and site B values (complex allowed). setup: hkl, lut_lo, lut_dim, mlambda, dynupz,
    rmt (3x3), sites [(Z, orbit[n,3])]. beams: (strong, strongsg, weak, weaksg)."""
    strong, ssg, weak, wsg = beams
    hkl = np.asarray(setup["hkl"], np.int64)
    lo, dims = setup["lut_lo"], setup["lut_dim"]
    lmb = float(setup["mlambda"])
    gs = hkl[np.asarray(strong)]
    gw = hkl[np.asarray(weak)] if len(weak) else np.zeros((0, 3), np.int64)
    n = len(strong)
    dt = np.result_type(lut_re, lut_im, np.asarray(site_b), np.float64)

    def U(d):
        cell, ok = _lut_lookup(lut_re, lo, dims, d)
        assert ok.all(), "beam pair outside the LUT box"
        return lut_re[cell], lut_im[cell]

    # DynMat in real-pair form
    Dr = np.zeros((n, n), dt)
    Di = np.zeros((n, n), dt)
    ur, ui = U(gs[:, None, :] - gs[None, :, :])
    if len(gw):
        ar, ai = U(gs[:, None, :] - gw[None, :, :])      # U(g - w)   [n, w]
        br, bi = U(gw[:, None, :] - gs[None, :, :])      # U(w - h)   [w, n]
        inv = 1.0 / np.asarray(wsg, np.float64)
        wr = (ar * inv) @ br - (ai * inv) @ bi
        wi = (ar * inv) @ bi + (ai * inv) @ br
        dsum = ((ar * ar + ai * ai) * inv).sum(1)
    else:
        wr = wi = np.zeros((n, n), dt)
        dsum = np.zeros(n, dt)
    c = 0.5 * lmb
    Dr[:] = ur - c * wr
    Di[:] = ui - c * wi
    idx = np.arange(n)
    Dr[idx, idx] = 2.0 * np.asarray(ssg, np.float64) / lmb - dsum * lmb / 2.0
    Di[idx, idx] = float(setup["dynupz"])
    # A = 2 pi i dz D / fl32(2kn) in the real 2n x 2n form
    pre = 2.0 * np.pi * dz / float(np.float32(2.0 * kn))
    Ar, Ai = -pre * Di, pre * Dr
    M = np.block([[Ar, -Ai], [Ai, Ar]])
    E = scipy.linalg.expm(M)
    psi = np.zeros(2 * n, dt)
    psi[0] = 1.0
    Lr = np.zeros((n, n), dt)
    Li = np.zeros((n, n), dt)
    for k in range(len(lam)):
        psi = E @ psi
        pr, pi = psi[:n], psi[n:]
        w = float(np.float32(lam[k]))
        Lr += w * (np.outer(pr, pr) + np.outer(pi, pi))
        Li += w * (np.outer(pi, pr) - np.outer(pr, pi))
    Lr *= dz / thick
    Li *= dz / thick
    # smooth S_p and the contraction
    rmt = np.asarray(setup["rmt"], np.float64).reshape(3, 3)
    d = gs[:, None, :] - gs[None, :, :]
    df = d.astype(np.float32).astype(np.float64)
    ln = np.sqrt(np.einsum("abi,ij,abj->ab", df, rmt, df).astype(np.float32))
    kkl = (np.float32(0.25) * (ln * ln).astype(np.float32)).astype(np.float64)
    out = []
    for p, (Z, orbit) in enumerate(setup["sites"]):
        z2 = float(np.float32(Z)) ** 2
        arg = 2.0 * np.pi * np.einsum("abi,ki->abk", d.astype(np.float64), np.asarray(orbit, np.float64))
        dw = z2 * np.exp(-site_b[p] * kkl)
        Sr = dw * np.cos(arg).sum(-1)
        Si = dw * np.sin(arg).sum(-1)
        out.append(np.sum(Lr * Sr - Li * Si))
    return np.array(out)


def derivative_lut(setup, beams, kn, lam, dz, thick, lut, site_b, cell, part):
    """d dots / d Re U(cell) (part 're') or d Im U(cell) (part 'im') by the complex step."""
    lr = np.asarray(lut.real, np.complex128).ravel().copy()
    li = np.asarray(lut.imag, np.complex128).ravel().copy()
    if part == "re":
        lr[cell] += 1j * H
    else:
        li[cell] += 1j * H
    b = np.asarray(site_b, np.complex128)
    return dots_forward(setup, beams, kn, lam, dz, thick, lr, li, b).imag / H


def derivative_b(setup, beams, kn, lam, dz, thick, lut, site_b, p):
    """d dots / d B_p by the complex step."""
    lr = np.asarray(lut.real, np.float64).ravel()
    li = np.asarray(lut.imag, np.float64).ravel()
    b = np.asarray(site_b, np.complex128).copy()
    b[p] += 1j * H
    return dots_forward(setup, beams, kn, lam, dz, thick, lr, li, b).imag / H


def setup_from_fixture(fx) -> dict:
    lut = np.asarray(fx["lut"]).astype(np.complex64).astype(np.complex128)  # the device's float32 LUT
    return dict(hkl=np.asarray(fx["hkl"], np.int64), lut_lo=np.asarray(fx["lut_lo"], np.int64),
                lut_dim=np.array(lut.shape[::-1], np.int64), mlambda=float(fx["mLambda"]),
                dynupz=float(np.float32(fx["DynUpz"])), rmt=np.asarray(fx["rmt"]), lut=lut,
                sites=[(int(fx[f"site{i}_Z"]), np.asarray(fx[f"site{i}_orbit"])) for i in range(int(fx["nsite"]))],
                site_b=np.array([float(np.float32(fx[f"site{i}_pos"][4])) for i in range(int(fx["nsite"]))]))
