"""Master-pattern residual and gradient (ctemsoft_amd.fit) for the adjoint workload's model:
fcc Ni, target = the same model with B x 1.05 and LUT x (1 + 0.01 N(0,1)), on the workload's own
Lambert grid (npx 100, 5,151 directions) and on a reduced one (npx 20) for the descent step."""
import sys

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _problem(npx=20):
    from ctemsoft_amd import ebsd, fit, workloads
    fx = load_golden("ni_c0w")
    st = ebsd.Setup.from_fixture(fx)
    wl = workloads.get("ni_c4w")
    ij, k, kn = ebsd.kvectors(npx, npx, st.isym, st.usehex, st.dsm, st.mlambda, 1)
    return fx, st, wl, ij, k, kn, fit.perturbed_target_setup(st, wl.seed)


@pytest.mark.parametrize("npx,precision", [(20, "fp64"), (100, "fp64"), (100, "mixed")])
def test_master_gradient_directional_and_descent(npx, precision):
    """The residual's gradient against a central difference of the FP64 loss along a random
    direction of every LUT entry; precision "mixed" takes loss and gradient from the mixed forward
    and reverse pass (the workload's mixed form), checked against the same FP64 difference."""
    from ctemsoft_amd import ebsd, fit
    fx, st, wl, ij, k, kn, st_t = _problem(npx)
    npx = int(ij[:, 0].max())
    isym = int(st.isym)
    ctx = ebsd.Context(0)
    lam = wl.depth_weights()
    thick = float(int(wl.nz * wl.dz))
    ctx.upload_crystal(st_t)
    pt = ctx.plan(k, kn, lam, wl.dz, thick)
    pt.run("fp64")
    target = pt.scatter(ij, npx, npx, isym)
    pt.close()
    ctx.upload_crystal(st)
    plan = ctx.plan(k, kn, lam, wl.dz, thick)
    owner = fit.pixel_owner(ij, npx, npx, isym)
    import time
    t0 = time.perf_counter()
    loss, glut, gb = fit.master_loss_grad(plan, ij, npx, npx, isym, target, owner, precision)
    t1 = time.perf_counter()
    _say(f"{precision}: {len(k)} directions, loss {loss:.6e}, |grad LUT| {np.abs(glut).max():.3e}, grad B {gb}, "
         f"forward + reverse {t1 - t0:.2f} s")
    assert loss > 0 and np.isfinite(glut).all() and np.isfinite(gb).all()
    # the loss from the mixed forward (the inner loop of a fit) against the FP64 one: the
    # residuals are ~1e-2 of the intensities, the mixed form's error ~1e-5 of them
    loss_m, _ = fit.master_loss(plan, ij, npx, npx, isym, target, precision="mixed")
    loss_d, _ = fit.master_loss(plan, ij, npx, npx, isym, target, precision="fp64")
    _say(f"loss fp64 {loss_d:.8e} mixed {loss_m:.8e}")
    assert abs(loss_m - loss_d) <= 1e-2 * loss_d
    plan.close()

    d = st.lut_dim
    lut0 = st.lut.view(np.complex128).reshape(d[2], d[1], d[0]).astype(np.complex64)
    rng = np.random.default_rng(9)
    v = (rng.standard_normal(lut0.shape) + 1j * rng.standard_normal(lut0.shape)) * np.abs(lut0)

    def loss_at(lut):
        s2 = ebsd.Setup(st.hkl, st.ucg.view(np.complex128), st.dbdiff, lut, st.lut_lo, st.mlambda,
                        st.dynupz, st.weakcutoff, st.cutoff, st.sgcutoff, st.rmt, st.sites, dsm=st.dsm,
                        isym=st.isym, usehex=st.usehex)
        ctx.upload_crystal(s2)
        p2 = ctx.plan(k, kn, lam, wl.dz, thick)
        val, _ = fit.master_loss(p2, ij, npx, npx, isym, target)
        p2.close()
        return val

    # the residual loss is quadratic in the intensities, so the central difference carries an
    # eps^3 sum(d1 d2) term that is not proportional to the residual: keep eps small
    eps = 2e-4
    lp = (lut0 + eps * v).astype(np.complex64)
    lm = (lut0 - eps * v).astype(np.complex64)
    delta = (lp.astype(np.complex128) - lm.astype(np.complex128))
    fd = loss_at(lp) - loss_at(lm)
    pred = float(np.sum(glut.real * delta.real + glut.imag * delta.imag))
    _say(f"directional: fd {fd:.6e} adjoint {pred:.6e}")
    assert abs(fd - pred) <= 2e-3 * abs(pred)

    if npx != 20:
        ctx.close()
        return
    # a step against the gradient lowers the residual
    g = glut.astype(np.complex128)
    step = 0.2 * loss / float(np.sum(np.abs(g) ** 2))
    l1 = loss_at((lut0 - step * g).astype(np.complex64))
    _say(f"descent: {loss:.6e} -> {l1:.6e}")
    assert l1 < loss
    ctx.close()
