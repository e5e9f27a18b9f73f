"""Master-pattern residual and its gradient with respect to the LUT (complex Ucg) and the site
Debye-Waller factors, on top of the C-ABI (ctem_plan_run / ctem_plan_scatter /
ctem_plan_adjoint, ctem_plan_adjoint_mixed).

    loss = sum over sites and Lambert pixels of (sr - sr_target)^2

sr is the float32 master array the reference program writes (CTEMEBSDmaster.f90:456-479):
svals = fl32(dots) / fl32(sum of orbit sizes) scattered to the Laue-equivalent pixels of each
direction, later directions overwriting earlier ones where equivalents coincide. The gradient
treats the float32 roundings as identities and holds the beam sets fixed.

The target of the adjoint workload is the same model with B scaled by 1.05 and every LUT entry
multiplied by (1 + 0.01 N(0,1)), one seeded factor shared by g and -g.
"""
from __future__ import annotations

import numpy as np

from . import ebsd


def pixel_owner(ij, npx, npy, isym):
    """For every pixel of sr[2npy+1][2npx+1], the index of the direction whose value the scatter
    leaves there (the last writer), or -1."""
    owner = -np.ones((2 * npy + 1) * (2 * npx + 1), np.int64)
    for d, (i, j) in enumerate(np.asarray(ij)):
        q = ebsd.laue_equiv(int(i), int(j), int(isym))
        q = q[(np.abs(q[:, 0]) <= npx) & (np.abs(q[:, 1]) <= npy)]
        owner[(q[:, 1] + npy) * (2 * npx + 1) + (q[:, 0] + npx)] = d
    return owner


def master_loss(plan, ij, npx, npy, isym, target, precision="fp64"):
    """Forward pass: (loss, sr)."""
    plan.run(precision)
    sr = plan.scatter(ij, npx, npy, isym)
    r = sr.astype(np.float64) - np.asarray(target, np.float64)
    return float(np.sum(r * r)), sr


def master_loss_grad(plan, ij, npx, npy, isym, target, owner=None, precision="fp64"):
    """(loss, dloss/dLUT shaped like the LUT [l][k][h], dloss/dB per site). precision "fp64":
    FP64 forward and reverse pass; "mixed": the loss from the mixed forward and the gradient from
    the mixed reverse pass (ctem_plan_adjoint_mixed, which reruns the same mixed forward)."""
    loss, sr = master_loss(plan, ij, npx, npy, isym, target, precision)
    if owner is None:
        owner = pixel_owner(ij, npx, npy, isym)
    nsite = plan.nsite
    r = (sr.astype(np.float64) - np.asarray(target, np.float64)).reshape(nsite, -1)
    nat = float(np.float32(plan.setup.nat))
    gsv = np.zeros((plan.numk, nsite))
    m = owner >= 0
    for p in range(nsite):
        np.add.at(gsv[:, p], owner[m], 2.0 * r[p, m])
    glut, gb = plan.adjoint(gsv / nat, precision)
    return loss, glut, gb


def perturbed_target_setup(setup: ebsd.Setup, seed: int, b_scale=1.05, u_sigma=0.01):
    """The adjoint workload's target model: B * b_scale, LUT(g) * (1 + u_sigma N(0,1)) with one
    factor per {g, -g} pair."""
    rng = np.random.default_rng(seed)
    d = setup.lut_dim
    lut = setup.lut.view(np.complex128).reshape(d[2], d[1], d[0])
    fac = 1.0 + u_sigma * rng.standard_normal(lut.shape)
    # the box is symmetric about g = 0 (lo = -hi), so -g is the reversed index along each axis
    fac = np.where(np.arange(fac.size).reshape(fac.shape) <= fac.size // 2, fac, fac[::-1, ::-1, ::-1])
    lut_t = (lut * fac).astype(np.complex64)
    st = ebsd.Setup(setup.hkl, setup.ucg.view(np.complex128), setup.dbdiff, lut_t, setup.lut_lo,
                    setup.mlambda, setup.dynupz, setup.weakcutoff, setup.cutoff, setup.sgcutoff,
                    setup.rmt, setup.sites, dsm=setup.dsm, isym=setup.isym, usehex=setup.usehex)
    st.site_b = (setup.site_b * np.float32(b_scale)).astype(np.float32)
    return st
