"""Python host mirror of the EBSD master-pattern path, over the C-ABI of
include/ctem_ebsd.h (ctypes). It plays the role of CTEMEBSDmaster's ComputeMasterPattern
(src_programs/CTEMEBSDmaster.f90:99-587) around the GPU beam loop:

    ctx = Context()                      # one GPU
    ctx.upload_crystal(setup)            # per-energy setup (LUT, reflection list, sites)
    ij, k, kn = kvectors(...)            # Calckvectors 'RoscaLambert'
    plan = ctx.plan(k, kn, lam, dz, thick)
    plan.run("fp64")                     # beam selection, DynMat, Lgh, Sgh contraction
    svals = plan.svals()                 # float32, as the reference stores them
    sr = plan.scatter(ij, npx, npy, isym)

Errors from the library raise CtemError with the library's message; there is no CPU path:
if the HIP library or the GPU is missing the calls fail.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libctem_ebsd.so")

OK = 0
PREC = {"fp64": 0, "mixed": 1}

EXPORTS = ["ctem_ctx_create", "ctem_ctx_destroy", "ctem_last_error", "ctem_device_info",
           "ctem_upload_crystal", "ctem_kvectors", "ctem_laue_equiv", "ctem_plan_create",
           "ctem_plan_destroy", "ctem_plan_select", "ctem_plan_get_beams", "ctem_plan_run",
           "ctem_plan_get_dots", "ctem_plan_get_svals", "ctem_plan_get_lgh", "ctem_plan_dynmat",
           "ctem_plan_scatter", "ctem_plan_timing", "ctem_calclgh3", "ctem_plan_set_mixed_terms",
           "ctem_plan_set_symmetry", "ctem_plan_adjoint", "ctem_plan_adjoint_mixed", "ctem_plan_alg_flops", "ctem_calclgh3_static",
           "ctem_plan_set_graph", "ctem_plans_run", "ctem_plan_set_tuning"]


class CtemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ctem error {code}: {msg}")
        self.code = code


class _Crystal(C.Structure):
    _fields_ = [("nref", C.c_int32), ("hkl", C.c_void_p), ("ucg", C.c_void_p),
                ("dbdiff", C.c_void_p), ("lut_lo", C.c_int32 * 3), ("lut_dim", C.c_int32 * 3),
                ("lut", C.c_void_p), ("mlambda", C.c_double), ("dynupz", C.c_float),
                ("weakcutoff", C.c_float), ("cutoff", C.c_float), ("sgcutoff", C.c_float),
                ("rmt", C.c_double * 9), ("nsite", C.c_int32), ("site_z", C.c_void_p),
                ("site_b", C.c_void_p), ("site_norbit", C.c_void_p), ("orbit", C.c_void_p)]


_LIB = None


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is missing: build it with `python -m ctemsoft_amd.build`")
        L = C.CDLL(LIBPATH)
        P, I32, D = C.c_void_p, C.c_int32, C.c_double
        L.ctem_last_error.restype = C.c_char_p
        sig = {
            "ctem_ctx_create": [I32, P], "ctem_ctx_destroy": [P],
            "ctem_device_info": [P, C.c_char_p, I32, P], "ctem_upload_crystal": [P, P],
            "ctem_kvectors": [I32, I32, I32, I32, I32, P, D, I32, P, P, P, P],
            "ctem_laue_equiv": [I32, I32, I32, P, P],
            "ctem_plan_create": [P, I32, P, P, I32, P, D, D, P], "ctem_plan_destroy": [P],
            "ctem_plan_select": [P, P, P], "ctem_plan_get_beams": [P, I32, P, P, P, P],
            "ctem_plan_run": [P, I32, I32], "ctem_plan_get_dots": [P, P],
            "ctem_plan_get_svals": [P, P], "ctem_plan_get_lgh": [P, P, I32, P],
            "ctem_plan_dynmat": [P, I32, P, I32, P], "ctem_plan_scatter": [P, P, I32, I32, I32, P],
            "ctem_plan_timing": [P, P, P, P, P], "ctem_plan_set_mixed_terms": [P, I32, I32],
            "ctem_plan_set_symmetry": [P, I32, P],
            "ctem_calclgh3": [P, P, P, D, D, I32, I32, D, P, I32],
            "ctem_calclgh3_static": [P, P, P, D, D, I32, I32, D, P, I32],
            "ctem_plan_adjoint": [P, P, P, P], "ctem_plan_adjoint_mixed": [P, P, P, P],
            "ctem_plan_alg_flops": [P, P, P, P],
            "ctem_plan_set_graph": [P, I32], "ctem_plans_run": [P, I32, I32],
            "ctem_plan_set_tuning": [P, I32, D],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = I32
        _LIB = L
    return _LIB


def _check(rc):
    if rc != OK:
        raise CtemError(rc, lib().ctem_last_error().decode())


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def lambert_dims(npx, usehex):
    """(npx, npy) of the Lambert sampling grid as the reference program sets them up
    (CTEMEBSDmaster.f90:270-271, 388-400): npy = npx for the square grid; for the hexagonal
    grid npy = nint(2 npx / sqrt(3)) evaluated in float32."""
    if usehex:
        v = np.float32(2.0) * np.float32(npx) / np.sqrt(np.float32(3.0))
        return int(npx), int(np.floor(float(v) + 0.5))
    return int(npx), int(npx)


# First space-group number of each of the 32 crystallographic point groups in International
# Tables order, and the Laue class (1..11: -1, 2/m, mmm, 4/m, 4/mmm, -3, -3m, 6/m, 6/mmm, m-3,
# m-3m) of each point group (symmetryvars.f90:242-244,286-287).
_SG_FIRST_OF_PG = (1, 2, 3, 6, 10, 16, 25, 47, 75, 81, 83, 89, 99, 111, 123, 143, 147, 149, 156, 162,
                   168, 174, 175, 177, 183, 187, 191, 195, 200, 207, 215, 221)
_LAUE_OF_PG = (1, 1, 2, 2, 2, 3, 3, 3, 4, 4, 4, 5, 5, 5, 5, 6, 6, 7, 7, 7, 8, 8, 8, 9, 9, 9, 9, 10, 10,
               11, 11, 11)
# trigonal -3m space groups whose Lambert sampling takes the second setting of the Laue group
# (isym 12; CTEMEBSDmaster.f90:254-282)
_LAUE_SECOND_SETTING = (149, 151, 153, 156, 158, 160, 161, 164, 165, 166, 167)


def run_plans(plans, precision="fp64"):
    """All plans (of one context) as one captured HIP graph (ctem_plans_run)."""
    arr = (C.c_void_p * len(plans))(*[p._h.value for p in plans])
    _check(lib().ctem_plans_run(C.cast(arr, C.c_void_p), len(plans), PREC[precision]))


def laue_group(sgnum):
    """(isym, usehex) that the master-pattern program derives from the space-group number
    (CTEMEBSDmaster.f90:249-282): the Laue class of the space group's point group (the last point
    group whose first space group is <= sgnum), 12 in place of 7 for the listed trigonal groups,
    and the hexagonal Lambert grid for Laue classes 6..9 and 12."""
    sgnum = int(sgnum)
    if not 1 <= sgnum <= 230:
        raise ValueError(f"space group {sgnum} not in 1..230")
    pg = max(i for i, first in enumerate(_SG_FIRST_OF_PG) if first <= sgnum)
    isym = _LAUE_OF_PG[pg]
    if isym == 7 and sgnum in _LAUE_SECOND_SETTING:
        isym = 12
    return isym, int(6 <= isym <= 9 or isym == 12)


def kvectors(npx, npy, isym, usehex, dsm, mlambda, hemisphere=1):
    """Incident directions of Calckvectors 'RoscaLambert' (host computation in the library)."""
    L = lib()
    dsm = np.ascontiguousarray(dsm, np.float64).ravel()
    n = C.c_int32()
    _check(L.ctem_kvectors(npx, npy, isym, int(usehex), hemisphere, _p(dsm), mlambda, 0, None,
                           None, None, C.byref(n)))
    cap = n.value
    ij = np.zeros(2 * cap, np.int32); k = np.zeros(3 * cap); kn = np.zeros(cap)
    _check(L.ctem_kvectors(npx, npy, isym, int(usehex), hemisphere, _p(dsm), mlambda, cap, _p(ij),
                           _p(k), _p(kn), C.byref(n)))
    return ij.reshape(-1, 2), k.reshape(-1, 3), kn


def laue_equiv(ipx, ipy, isym):
    q = np.zeros(24, np.int32); n = C.c_int32()
    _check(lib().ctem_laue_equiv(ipx, ipy, isym, _p(q), C.byref(n)))
    return q[:2 * n.value].reshape(-1, 2)


class Setup:
    """Per-energy crystal setup arrays in the layout of ctem_crystal."""

    def __init__(self, hkl, ucg, dbdiff_ref, lut, lut_lo, mlambda, dynupz, weakcutoff, cutoff,
                 sgcutoff, rmt, sites, dsm=None, isym=None, usehex=0):
        self.hkl = np.ascontiguousarray(hkl, np.int32).reshape(-1, 3)
        u = np.asarray(ucg, np.complex128)
        self.ucg = np.ascontiguousarray(np.stack([u.real, u.imag], -1).ravel())
        self.dbdiff = np.ascontiguousarray(dbdiff_ref, np.uint8)
        lutc = np.asarray(lut, np.complex128)  # indexed [l][k][h]
        self.lut_dim = np.array(lutc.shape[::-1], np.int32)
        self.lut_lo = np.asarray(lut_lo, np.int32)
        self.lut = np.ascontiguousarray(np.stack([lutc.real, lutc.imag], -1).ravel())
        self.mlambda, self.dynupz = float(mlambda), float(dynupz)
        self.weakcutoff, self.cutoff, self.sgcutoff = float(weakcutoff), float(cutoff), float(sgcutoff)
        self.rmt = np.ascontiguousarray(rmt, np.float64).ravel()
        self.sites = sites
        self.site_z = np.array([s["Z"] for s in sites], np.int32)
        self.site_b = np.array([s["B"] for s in sites], np.float32)
        self.site_norbit = np.array([len(s["orbit"]) for s in sites], np.int32)
        self.orbit = np.ascontiguousarray(np.concatenate([np.asarray(s["orbit"], np.float64).ravel()
                                                          for s in sites]))
        self.nat = int(self.site_norbit.sum())
        self.dsm = None if dsm is None else np.ascontiguousarray(dsm, np.float64).ravel()
        self.isym, self.usehex = isym, usehex

    @classmethod
    def from_fixture(cls, fx, lut_from=None):
        """Setup arrays from a fixture (.npz). With lut_from (a fixture of the same crystal and
        energy that holds `lut_full`), the LUT is the one over the box that holds every
        difference of two list entries instead of the reference-sized one."""
        lut = np.asarray(fx["lut"])
        lo = np.asarray(fx["lut_lo"])
        hkl = np.asarray(fx["hkl"], np.int32)
        h, k, l = (hkl - lo).T
        dbd = np.asarray(fx["dbdiff"])[l, k, h]
        if lut_from is not None:
            if not np.array_equal(np.asarray(lut_from["hkl"]), np.asarray(fx["hkl"])):
                raise ValueError("lut_from holds another reflection list")
            lut = np.asarray(lut_from["lut_full"])
            lo = np.asarray(lut_from["lut_full_lo"])
        sites = [dict(Z=int(fx[f"site{i}_Z"]), B=float(fx[f"site{i}_pos"][4]),
                      orbit=np.asarray(fx[f"site{i}_orbit"])) for i in range(int(fx["nsite"]))]
        wc, c, sgc = (float(v) for v in fx["cutoffs"])
        return cls(hkl, fx["ucg"], dbd, lut, lo, float(fx["mLambda"]), float(fx["DynUpz"]), wc, c,
                   sgc, fx["rmt"], sites, dsm=fx["dsm"], isym=int(fx["isym"]), usehex=int(fx["usehex"]))

    def struct(self):
        s = _Crystal()
        s.nref = len(self.hkl)
        s.hkl, s.ucg, s.dbdiff, s.lut = _p(self.hkl), _p(self.ucg), _p(self.dbdiff), _p(self.lut)
        for i in range(3):
            s.lut_lo[i] = int(self.lut_lo[i]); s.lut_dim[i] = int(self.lut_dim[i])
        s.mlambda, s.dynupz = self.mlambda, self.dynupz
        s.weakcutoff, s.cutoff, s.sgcutoff = self.weakcutoff, self.cutoff, self.sgcutoff
        for i in range(9):
            s.rmt[i] = float(self.rmt[i])
        s.nsite = len(self.sites)
        s.site_z, s.site_b = _p(self.site_z), _p(self.site_b)
        s.site_norbit, s.orbit = _p(self.site_norbit), _p(self.orbit)
        return s


class Context:
    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().ctem_ctx_create(device, C.byref(self._h)))
        self.setup = None
        self._plans = weakref.WeakSet()

    def close(self):
        if self._h:
            for p in list(self._plans):
                p.close()
            lib().ctem_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self):
        buf = C.create_string_buffer(256); cu = C.c_int32()
        _check(lib().ctem_device_info(self._h, buf, 256, C.byref(cu)))
        return buf.value.decode(), cu.value

    def upload_crystal(self, setup: Setup):
        st = setup.struct()
        _check(lib().ctem_upload_crystal(self._h, C.byref(st)))
        self.setup = setup

    def plan(self, k, kn, lam, dz, thick):
        p = Plan(self, k, kn, lam, dz, thick)
        self._plans.add(p)
        return p

    def calclgh3(self, dmat, thick, kn, depthstep, lam):
        """Drop-in for CalcLgh3 (CTEMEBSDmaster.f90:610-722): DynMat -> Lgh, FP64 on the GPU."""
        d = np.asfortranarray(dmat, np.complex128)
        n = d.shape[0]
        out = np.zeros((n, n), np.complex128, order="F")
        lam = np.ascontiguousarray(lam, np.float32)
        _check(lib().ctem_calclgh3(self._h, d.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
                                   float(thick), float(kn), n, 1, float(depthstep), _p(lam), len(lam)))
        return np.array(out)


class Plan:
    def __init__(self, ctx: Context, k, kn, lam, dz, thick):
        self.ctx = ctx
        self.k = np.ascontiguousarray(k, np.float64).reshape(-1, 3)
        self.kn = np.ascontiguousarray(kn, np.float64)
        self.lam = np.ascontiguousarray(lam, np.float32)
        self.numk = len(self.k)
        # the crystal current at creation (the library keeps it with the plan)
        self.setup = ctx.setup
        self.nsite = len(ctx.setup.sites) if ctx.setup is not None else 0
        self._h = C.c_void_p()
        _check(lib().ctem_plan_create(ctx._h, self.numk, _p(self.k), _p(self.kn), len(self.lam),
                                      _p(self.lam), float(dz), float(thick), C.byref(self._h)))
        self.nns = self.nnw = None

    def close(self):
        if self._h and self.ctx._h:
            lib().ctem_plan_destroy(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def select(self):
        ns = np.zeros(self.numk, np.int32); nw = np.zeros(self.numk, np.int32)
        _check(lib().ctem_plan_select(self._h, _p(ns), _p(nw)))
        self.nns, self.nnw = ns, nw
        return ns, nw

    def beams(self, ik):
        ns, nw = int(self.nns[ik]), int(self.nnw[ik])
        s = np.zeros(ns, np.int32); ssg = np.zeros(ns, np.float32)
        w = np.zeros(max(nw, 1), np.int32); wsg = np.zeros(max(nw, 1), np.float32)
        _check(lib().ctem_plan_get_beams(self._h, ik, _p(s), _p(ssg), _p(w), _p(wsg)))
        return s, ssg, w[:nw], wsg[:nw]

    def set_mixed_terms(self, taylor=16, square=16):
        """Split schemes of the mixed form: 3 = BF16x3, 6 = BF16x6, 16 = FP16x3 (default)."""
        _check(lib().ctem_plan_set_mixed_terms(self._h, int(taylor), int(square)))

    def set_symmetry(self, mode=1):
        """Symmetric products in the split mixed form: 1 = when the crystal's structure-factor
        table is inversion-symmetric (default), 0 = never. Returns whether they are in effect."""
        on = C.c_int32()
        _check(lib().ctem_plan_set_symmetry(self._h, int(mode), C.byref(on)))
        return bool(on.value)

    def run(self, precision="fp64", keep_lgh=-1):
        _check(lib().ctem_plan_run(self._h, PREC[precision], int(keep_lgh)))

    def set_tuning(self, bucket_multiple=32, chunk_gib=24.0):
        """Bucket granularity of the padded sizes and work space per chunk (ctem_plan_set_tuning)."""
        _check(lib().ctem_plan_set_tuning(self._h, int(bucket_multiple), float(chunk_gib)))

    def set_graph(self, on=True):
        """Capture the plan's run into a HIP graph once its squaring counts are known, and
        replay it (ctem_plan_set_graph)."""
        _check(lib().ctem_plan_set_graph(self._h, 1 if on else 0))

    def dots(self):
        d = np.zeros(self.numk * self.nsite)
        _check(lib().ctem_plan_get_dots(self._h, _p(d)))
        return d.reshape(self.numk, self.nsite)

    def adjoint(self, gdots, precision="fp64"):
        """Reverse mode: gdots[ik, p] = df/ddots -> (df/dU as a complex array shaped like the
        uploaded LUT, [l][k][h], with df/dRe U + i df/dIm U; df/dB per site). precision "fp64"
        (ctem_plan_adjoint) or "mixed" (ctem_plan_adjoint_mixed: the mixed forward, split-FP16
        products in the derivative of the exponential)."""
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', not {precision!r}")
        st = self.setup
        g = np.ascontiguousarray(gdots, np.float64).reshape(self.numk, self.nsite)
        ncell = int(np.prod(st.lut_dim))
        glut = np.zeros(2 * ncell)
        gb = np.zeros(max(1, self.nsite))
        f = lib().ctem_plan_adjoint if precision == "fp64" else lib().ctem_plan_adjoint_mixed
        _check(f(self._h, _p(g), _p(glut), _p(gb)))
        d = st.lut_dim
        return glut.view(np.complex128).reshape(d[2], d[1], d[0]), gb[:self.nsite].copy()

    def svals(self):
        s = np.zeros(self.numk * self.nsite, np.float32)
        _check(lib().ctem_plan_get_svals(self._h, _p(s)))
        return s.reshape(self.numk, self.nsite)

    def lgh(self):
        n = C.c_int32()
        _check(lib().ctem_plan_get_lgh(self._h, None, 0, C.byref(n)))
        out = np.zeros(2 * n.value * n.value)
        _check(lib().ctem_plan_get_lgh(self._h, _p(out), n.value * n.value, C.byref(n)))
        return out.view(np.complex128).reshape(n.value, n.value)

    def dynmat(self, ik):
        n = C.c_int32()
        _check(lib().ctem_plan_dynmat(self._h, ik, None, 0, C.byref(n)))
        out = np.zeros(2 * n.value * n.value)
        _check(lib().ctem_plan_dynmat(self._h, ik, _p(out), n.value * n.value, C.byref(n)))
        return out.view(np.complex128).reshape(n.value, n.value)

    def scatter(self, ij, npx, npy, isym, sr=None):
        if sr is None:
            sr = np.zeros((self.nsite, 2 * npy + 1, 2 * npx + 1), np.float32)
        ij = np.ascontiguousarray(ij, np.int32)
        _check(lib().ctem_plan_scatter(self._h, _p(ij), npx, npy, isym, _p(sr)))
        return sr

    def timing(self):
        t, g, f, n = C.c_double(), C.c_double(), C.c_double(), C.c_int32()
        _check(lib().ctem_plan_timing(self._h, C.byref(t), C.byref(g), C.byref(f), C.byref(n)))
        return dict(total_ms=t.value, gemm_ms=g.value, gemm_flops=f.value, gemm_launches=n.value)

    def alg_flops(self):
        """Algorithmic flops of the last FP64 run by the reference's Pade rule: total path,
        matrix-exponential products 8n^3(q+j), Pade solve (32/3)n^3."""
        t, q, v = C.c_double(), C.c_double(), C.c_double()
        _check(lib().ctem_plan_alg_flops(self._h, C.byref(t), C.byref(q), C.byref(v)))
        return dict(total=t.value, products=q.value, solve=v.value)
