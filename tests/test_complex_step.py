"""Pins of the complex-step reference (oracle/complex_step.py) that the reverse mode is checked
against: its forward equals the oracle's dots (to the float32 rounding of the reference's S
terms, which the complex-step forward keeps smooth), and its derivatives equal central
differences of the same forward in double (to the truncation and rounding of a difference
quotient)."""
import numpy as np

from conftest import load_golden


def _case(oracle_mod, name="ni_c0w", j=1):
    from oracle import complex_step as cs
    fx = load_golden(name)
    st = cs.setup_from_fixture(fx)
    cr = oracle_mod.Crystal(fx)
    k = int(fx["dyn_ik"][j])
    beams = oracle_mod.select_beams(cr, fx["kvec"][k])
    args = (st, beams, float(fx["kn"][k]), fx["lam"], float(fx["dz"]), float(fx["thick"]))
    return fx, cr, k, st, beams, args


def test_forward_matches_oracle(oracle_mod):
    from oracle import complex_step as cs
    fx, cr, k, st, beams, args = _case(oracle_mod)
    lut = st["lut"].ravel()
    d = cs.dots_forward(*args, lut.real.copy(), lut.imag.copy(), st["site_b"])
    ref, _, _ = oracle_mod.master_direction(cr, fx["kvec"][k], fx["kn"][k], fx["lam"], float(fx["dz"]),
                                            float(fx["thick"]))
    assert np.allclose(d, ref, rtol=2e-6, atol=0), (d, ref)


def test_complex_step_matches_central_differences(oracle_mod):
    from oracle import complex_step as cs
    fx, cr, k, st, beams, args = _case(oracle_mod)
    lut = st["lut"]
    flat = lut.ravel()
    hkl = st["hkl"]
    s = np.asarray(beams[0])
    # cells that this direction's matrix reads: g - h of two strong beams
    d = hkl[s[1]] - hkl[s[0]]
    lo, dims = st["lut_lo"], st["lut_dim"]
    cells = [int((d[0] - lo[0]) + dims[0] * ((d[1] - lo[1]) + dims[1] * (d[2] - lo[2]))), int(np.argmax(np.abs(flat)))]
    for cell in cells:
        for part in ("re", "im"):
            g = cs.derivative_lut(*args[:6], lut, st["site_b"], cell, part)
            h = 1e-4 * max(abs(flat[cell]), 1e-2)
            lr, li = flat.real.copy(), flat.imag.copy()
            (lr if part == "re" else li)[cell] += h
            fp = cs.dots_forward(*args, lr, li, st["site_b"])
            (lr if part == "re" else li)[cell] -= 2 * h
            fm = cs.dots_forward(*args, lr, li, st["site_b"])
            fd = (fp - fm) / (2 * h)
            assert np.allclose(g, fd, rtol=1e-5, atol=1e-7 * np.abs(fd).max() + 1e-12), (cell, part, g, fd)
    gb = cs.derivative_b(*args[:6], lut, st["site_b"], 0)
    h = 1e-6 * st["site_b"][0]
    bp, bm = st["site_b"].copy(), st["site_b"].copy()
    bp[0] += h
    bm[0] -= h
    fd = (cs.dots_forward(*args, flat.real, flat.imag, bp) - cs.dots_forward(*args, flat.real, flat.imag, bm)) / (2 * h)
    assert np.allclose(gb, fd, rtol=1e-6), (gb, fd)
