"""Reverse mode (ctem_plan_adjoint) against central finite differences of the FP64 forward
(itself pinned to the reference by test_gpu_parity.py). The LUT is float32 on the device, so
each perturbed entry is rounded to float32 and the exact step (a difference of two float32
values) is used in the quotient."""
import sys

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from ctemsoft_amd import ebsd
    c = ebsd.Context(0)
    yield c
    c.close()


def _setup(fx, lut=None, bscale=None):
    from ctemsoft_amd import ebsd
    st = ebsd.Setup.from_fixture(fx)
    if lut is not None:
        st = ebsd.Setup(st.hkl, np.asarray(fx["ucg"]), st.dbdiff, lut, st.lut_lo, st.mlambda, st.dynupz,
                        st.weakcutoff, st.cutoff, st.sgcutoff, st.rmt, st.sites, dsm=fx["dsm"],
                        isym=int(fx["isym"]), usehex=int(fx["usehex"]))
    if bscale is not None:
        st.site_b = (st.site_b * np.float32(bscale)).astype(np.float32)
    return st


def _dots(ctx, fx, ks, st):
    ctx.upload_crystal(st)
    plan = ctx.plan(fx["kvec"][ks], fx["kn"][ks], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    plan.run("fp64")
    d = plan.dots().copy()
    plan.close()
    return d


@pytest.mark.parametrize("name", ["ni_c0w", "fe_c1"])
def test_lut_gradient_matches_finite_differences(ctx, name):
    fx = load_golden(name)
    rng = np.random.default_rng(3)
    ks = np.sort(rng.choice(len(fx["kvec"]), 6, replace=False))
    st = _setup(fx)
    ctx.upload_crystal(st)
    plan = ctx.plan(fx["kvec"][ks], fx["kn"][ks], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    plan.run("fp64")
    w = rng.standard_normal(plan.dots().shape)
    glut, gb = plan.adjoint(w)
    plan.close()
    lut0 = np.asarray(fx["lut"]).astype(np.complex64)
    # the cells with the largest gradients, real and imaginary directions
    flat = np.argsort(-np.abs(glut).ravel())[:4]
    for cell in flat:
        idx = np.unravel_index(cell, lut0.shape)
        for part in (1.0, 1.0j):
            u = complex(lut0[idx])
            h = 2.0 ** -8 * max(abs(u), 1e-3)
            lp, lm = lut0.copy(), lut0.copy()
            lp[idx] = np.complex64(u + part * h)
            lm[idx] = np.complex64(u - part * h)
            step = (complex(lp[idx]) - complex(lm[idx])) / part
            fp = float(np.sum(w * _dots(ctx, fx, ks, _setup(fx, lut=lp))))
            fm = float(np.sum(w * _dots(ctx, fx, ks, _setup(fx, lut=lm))))
            fd = (fp - fm) / step.real
            ad = glut[idx].real if part == 1.0 else glut[idx].imag
            _say(f"{name} cell {idx} {'re' if part == 1.0 else 'im'}: adjoint {ad:.6e} fd {fd:.6e}")
            assert abs(ad - fd) <= 1e-4 * np.abs(glut).max() + 1e-3 * abs(fd), (idx, part, ad, fd)


def test_debye_waller_gradient_matches_direct_sum(ctx):
    """d dots_p / d B_p = Re sum_gh Lgh(g,h) dS_p(g-h)/dB_p, with dS_p/dB_p summed here over the
    orbit in double (-|g-h|^2/4 Z^2 exp(-B |g-h|^2/4) e^{2 pi i (g-h).r}) and Lgh from the GPU.
    A finite difference in B is no check here: the float32 terms of S make the forward
    piecewise constant in B at a level that the cancellation in Re sum Lgh S magnifies to
    ~1e-3 of the derivative."""
    from ctemsoft_amd import ebsd
    fx = load_golden("al2o3_c2")
    rng = np.random.default_rng(4)
    ks = np.sort(rng.choice(len(fx["kvec"]), 4, replace=False))
    st = _setup(fx)
    ctx.upload_crystal(st)
    plan = ctx.plan(fx["kvec"][ks], fx["kn"][ks], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    plan.select()
    w = rng.standard_normal((len(ks), 2))
    plan.run("fp64")
    _, gb = plan.adjoint(w)
    rmt = np.asarray(fx["rmt"]).reshape(3, 3)
    hkl = fx["hkl"].astype(np.int64)
    ref = np.zeros(2)
    for j in range(len(ks)):
        plan.run("fp64", keep_lgh=j)
        L = plan.lgh()
        g = hkl[plan.beams(j)[0]]
        d = g[:, None, :] - g[None, :, :]
        df = d.astype(np.float32).astype(np.float64)
        ln = np.sqrt(np.einsum("abi,ij,abj->ab", df, rmt, df).astype(np.float32))
        kkl = (np.float32(0.25) * (ln * ln).astype(np.float32)).astype(np.float64)
        for p in range(2):
            Z = float(np.float32(fx[f"site{p}_Z"])) ** 2
            B = float(np.float32(fx[f"site{p}_pos"][4]))
            ph = np.exp(2j * np.pi * np.einsum("abi,ki->abk", d.astype(np.float64), fx[f"site{p}_orbit"])).sum(-1)
            dS = -kkl * Z * np.exp(-B * kkl) * ph
            ref[p] += w[j, p] * float(np.real(np.sum(L * dS)))
    plan.close()
    _say(f"dB adjoint {gb} direct {ref}")
    assert np.allclose(gb, ref, rtol=1e-9, atol=0)


def test_directional_derivative_whole_lut(ctx):
    """<grad, v> for a random perturbation v of every LUT entry at once."""
    fx = load_golden("ni_c0w")
    rng = np.random.default_rng(5)
    ks = np.sort(rng.choice(len(fx["kvec"]), 16, replace=False))
    st = _setup(fx)
    ctx.upload_crystal(st)
    plan = ctx.plan(fx["kvec"][ks], fx["kn"][ks], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    plan.run("fp64")
    w = rng.standard_normal(plan.dots().shape)
    glut, _ = plan.adjoint(w)
    plan.close()
    lut0 = np.asarray(fx["lut"]).astype(np.complex64)
    eps = 1e-3
    v = (rng.standard_normal(lut0.shape) + 1j * rng.standard_normal(lut0.shape)) * np.abs(lut0)
    lp = (lut0 + eps * v).astype(np.complex64)
    lm = (lut0 - eps * v).astype(np.complex64)
    dp = lp.astype(np.complex128) - lut0
    dm = lut0 - lm.astype(np.complex128)
    fp = float(np.sum(w * _dots(ctx, fx, ks, _setup(fx, lut=lp))))
    fm = float(np.sum(w * _dots(ctx, fx, ks, _setup(fx, lut=lm))))
    # first order: f(+) - f(-) = Re <grad, dp + dm>
    pred = float(np.sum(glut.real * (dp + dm).real + glut.imag * (dp + dm).imag))
    _say(f"directional: fd {fp - fm:.6e} adjoint {pred:.6e}")
    assert abs((fp - fm) - pred) <= 1e-4 * abs(pred)


def test_gradient_matches_complex_step(ctx, oracle_mod):
    """The reverse mode against the complex-step derivative (oracle/complex_step.py: the same
    forward carried in complex arithmetic on the real parameters, exact to rounding) of
    sum_j w_j dots_j over a few fcc-Ni directions: the real and imaginary parts of the LUT
    entries with the largest gradients, and the site's Debye-Waller factor B. The beam sets
    are held fixed, as the reverse mode holds them."""
    from oracle import complex_step as cs
    fx = load_golden("ni_c0w")
    st = cs.setup_from_fixture(fx)
    cr = oracle_mod.Crystal(fx)
    rng = np.random.default_rng(11)
    ks = np.sort(rng.choice(len(fx["kvec"]), 3, replace=False))
    ctx.upload_crystal(_setup(fx))
    plan = ctx.plan(fx["kvec"][ks], fx["kn"][ks], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    plan.run("fp64")
    w = rng.standard_normal(plan.dots().shape)
    glut, gb = plan.adjoint(w)
    plan.close()
    cases = []
    for j, k in enumerate(ks):
        beams = oracle_mod.select_beams(cr, fx["kvec"][k])
        cases.append((beams, float(fx["kn"][k]), w[j, 0]))
    args = lambda c: (st, c[0], c[1], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    flat = np.argsort(-np.abs(glut).ravel())[:3]
    gmax = np.abs(glut).max()
    for cell in flat:
        idx = np.unravel_index(cell, glut.shape)
        for part in ("re", "im"):
            ref = sum(c[2] * cs.derivative_lut(*args(c), st["lut"], st["site_b"], int(cell), part)[0] for c in cases)
            ad = glut[idx].real if part == "re" else glut[idx].imag
            _say(f"cell {idx} {part}: adjoint {ad:.12e} complex step {ref:.12e}")
            # the device contracts with the reference's float32-rounded S terms, the complex step
            # with the smooth S they round: ~1e-7 relative between the two
            assert abs(ad - ref) <= 1e-7 * gmax + 1e-6 * abs(ref), (idx, part, ad, ref)
    refb = sum(c[2] * cs.derivative_b(*args(c), st["lut"], st["site_b"], 0)[0] for c in cases)
    _say(f"B: adjoint {gb[0]:.12e} complex step {refb:.12e}")
    assert abs(gb[0] - refb) <= 1e-8 * abs(refb)


# the mixed reverse pass (ctem_plan_adjoint_mixed) against the FP64 one on the same directions and
# weights: the plan's mixed forward in each of its forms (split-FP16 storage with symmetric
# products, with full products, and the split-BF16 options on FP32 storage), a bcc and a two-site
# trigonal crystal
MIXED_CASES = [("ni_c0w", 1, 16), ("ni_c0w", 0, 16), ("ni_c0w", 1, 3), ("fe_c1w", 1, 16), ("al2o3_c2w", 1, 16)]


@pytest.mark.parametrize("name,symmetry,terms", MIXED_CASES)
def test_mixed_adjoint_matches_fp64(ctx, name, symmetry, terms):
    from ctemsoft_amd import ebsd
    fx = load_golden(name)
    rng = np.random.default_rng(21)
    ks = np.sort(rng.choice(len(fx["kvec"]), 12, replace=False))
    # corundum: the full-box LUT (a reference-sized one does not hold every g-h of its list)
    st = ebsd.Setup.from_fixture(fx, load_golden("al2o3_c2") if name.startswith("al2o3") else None)
    ctx.upload_crystal(st)
    plan = ctx.plan(fx["kvec"][ks], fx["kn"][ks], fx["lam"], float(fx["dz"]), float(fx["thick"]))
    plan.set_symmetry(symmetry)
    plan.set_mixed_terms(terms, terms)
    plan.run("fp64")
    d64 = plan.dots().copy()
    w = rng.standard_normal(d64.shape)
    g64, b64 = plan.adjoint(w)
    plan.run("mixed")
    dmx = plan.dots().copy()
    gmx, bmx = plan.adjoint(w, "mixed")
    # the mixed reverse pass reruns the plan's mixed forward: the same intensities
    plan.run("mixed")
    assert np.array_equal(plan.dots(), dmx)
    plan.close()
    assert np.isfinite(gmx).all() and np.isfinite(bmx).all()
    eg = np.abs(gmx - g64).max() / np.abs(g64).max()
    eb = np.abs(bmx - b64).max() / np.abs(b64).max()
    ed = np.abs(dmx - d64).max() / np.abs(d64).max()
    _say(f"{name} symmetry {symmetry} terms {terms}: forward {ed:.2e}, grad LUT {eg:.2e}, grad B {eb:.2e}")
    tol = 2e-3 if terms == 3 else 2e-4
    assert eg <= tol and eb <= tol, (eg, eb)
