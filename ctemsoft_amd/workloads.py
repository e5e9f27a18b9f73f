"""Named EBSD master-pattern workloads (crystal, beam energy, Bethe cutoffs, Lambert grid,
depth profile), all seeded so that every run sees the same synthetic inputs.

Names follow the crystal and the role of the workload:

  ni_c0     fcc Ni (Fm-3m, a=0.35236 nm), 20 kV, dmin 0.1 nm, npx 50, one hemisphere,
            20 x 1 nm depth steps, w ~ exp(-z/15 nm)(1 + 0.05 N(0,1))
  fe_c1     bcc Fe (Im-3m, a=0.28665 nm), 20 kV, dmin 0.05 nm, npx 250, both hemispheres,
            50 x 0.5 nm depth steps, w ~ exp(-z/20 nm)
  al2o3_c2  corundum-type R-3c (hexagonal axes a=0.4759, c=1.2991 nm), 15 kV, dmin 0.05 nm,
            npx 200, both hemispheres, depth as fe_c1 (no depth profile is specified for it)
  fe_c3     bcc Fe energy series 10..20 kV in 0.5 kV steps, npx 150 per hemisphere,
            60 depth bins per energy from a gamma(k=2, theta=E/2 nm) histogram
  ni_c4     fcc Ni as ni_c0 with npx 100 (adjoint / gradient workload)

Bethe parameters are the reference's `&BetheList` namelist (gvectors.f90:1196-1234):
a reflection is strong when |s_g| <= weakcutoff * lambda|U_g| and weak when
|s_g| <= cutoff * lambda|U_g| (gvectors.f90:1025-1060). The plain names use the workload
definition's "4/8/50" as weakcutoff=4, cutoff=8, sgcutoff=0.05 (the third number has no
counterpart in this code). With that mapping the reference keeps only 1-12 strong beams per
direction, far below the 50-800 the workload definitions quote. The "w" variants
(ni_c0w, fe_c1w, al2o3_c2w, fe_c3w, ni_c4w) keep everything else and scale the two
cutoffs so that the strong-beam counts land in the quoted ranges:

  ni_c0w / ni_c4w   weakcutoff 240, cutoff 480    (quoted 50-150 / 100-250 strong beams)
  fe_c1w / fe_c3w   weakcutoff 320, cutoff 640    (quoted 80-300 / 60-350)
  al2o3_c2w         weakcutoff 320, cutoff 640    (quoted 200-800)
  al2o3_c2x         weakcutoff 640, cutoff 1280   (the upper half of 200-800: 509-616 strong beams)

Debye-Waller B is in nm^2, as in the reference's .xtal files (crystal.f90:1163).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Workload:
    name: str
    xtal: dict
    ekev: float
    dmin: float
    npx: int
    nz: int
    dz: float
    seed: int
    hemispheres: int = 1
    weakcutoff: float = 4.0
    cutoff: float = 8.0
    sgcutoff: float = 0.05
    profile: str = "exp"
    decay_nm: float = 15.0
    noise: float = 0.0
    energies: tuple = field(default_factory=tuple)

    def depth_weights(self, ekev: float | None = None) -> np.ndarray:
        """Normalised float32 depth weights lambda_k for k = 1..nz (z_k = k * dz)."""
        rng = np.random.default_rng(self.seed)
        if self.profile == "exp":
            z = self.dz * np.arange(1, self.nz + 1)
            w = np.exp(-z / self.decay_nm)
            if self.noise:
                w = w * (1.0 + self.noise * rng.standard_normal(self.nz))
        elif self.profile == "gamma":
            e = self.ekev if ekev is None else ekev
            theta = e / 2.0
            dz = theta / 5.0
            # per-energy seed so that each bin has its own histogram
            rng = np.random.default_rng(self.seed + int(round(e * 10)))
            samp = rng.gamma(2.0, theta, 1_000_000)
            w, _ = np.histogram(samp, bins=self.nz, range=(0.0, self.nz * dz))
            w = w.astype(np.float64)
        else:
            raise ValueError(self.profile)
        w = w / w.sum()
        return w.astype(np.float32)

    def describe(self) -> str:
        """One line naming the crystal, energy, grid, Bethe cutoffs and depth sampling."""
        x = self.xtal
        sg = {225: "Fm-3m", 229: "Im-3m", 167: "R-3c"}.get(x["sg"], f"SG {x['sg']}")
        cell = (f"a={x['abc'][0]:g} nm" if x["system"] == 1 else
                f"a={x['abc'][0]:g} c={x['abc'][2]:g} nm (hexagonal axes)")
        sites = "+".join(f"Z{st['Z']}" for st in x["sites"])
        energy = (f"{self.energies[0]:g}..{self.energies[-1]:g} kV ({len(self.energies)} bins)"
                  if self.energies else f"{self.ekev:g} kV")
        depth = (f"{self.nz} depth bins of E/10 nm" if self.profile == "gamma" else
                 f"{self.nz} x {self.dz:g} nm depth steps")
        return (f"{sg} {cell}, sites {sites}, {energy}, dmin {self.dmin:g} nm, Bethe "
                f"{self.weakcutoff:g}/{self.cutoff:g}/{self.sgcutoff:g}, npx {self.npx}, "
                f"{self.hemispheres} hemisphere{'s' if self.hemispheres > 1 else ''}, {depth}")

    def dz_for(self, ekev: float | None = None) -> float:
        if self.profile == "gamma":
            e = self.ekev if ekev is None else ekev
            return e / 2.0 / 5.0
        return self.dz

    def depth_weights_for_program(self, ekev: float | None = None):
        """Weights as the reference program would see them through its Monte Carlo file:
        integer counts per depth bin, lambda = f32(f32(count) / num_el)
        (CTEMEBSDmaster.f90:309-315), plus the tail count that puts the 95% depth cut on
        the last bin."""
        w = self.depth_weights(ekev).astype(np.float64)
        counts = np.rint(w * 1.0e9).astype(np.int64)
        S = int(counts.sum())
        tail = int(S * 0.05 / 0.95 * 0.99)
        cum = np.cumsum(counts)
        if not cum[-2] < 0.95 * (S + tail):
            # a profile whose last bin holds almost nothing (gamma): the largest tail that
            # still keeps the cut on the last bin
            tail = S // 19 - 1
        assert cum[-2] < 0.95 * (S + tail) <= cum[-1]
        lam = (counts.astype(np.float32).astype(np.float64) / float(S)).astype(np.float32)
        return lam, counts, tail


def _ni():
    return dict(system=1, abc=(0.35236, 0.35236, 0.35236), angles=(90.0, 90.0, 90.0), sg=225,
                sgset=1, sites=[dict(Z=28, xyz=(0.0, 0.0, 0.0), occ=1.0, B=0.0035)])


def _fe():
    return dict(system=1, abc=(0.28665, 0.28665, 0.28665), angles=(90.0, 90.0, 90.0), sg=229,
                sgset=1, sites=[dict(Z=26, xyz=(0.0, 0.0, 0.0), occ=1.0, B=0.0030)])


def _al2o3():
    return dict(system=4, abc=(0.4759, 0.4759, 1.2991), angles=(90.0, 90.0, 120.0), sg=167,
                sgset=1, sites=[dict(Z=13, xyz=(0.0, 0.0, 0.3522), occ=1.0, B=0.003),
                                dict(Z=8, xyz=(0.3064, 0.0, 0.25), occ=1.0, B=0.003)])


_WORKLOADS = {
    "ni_c0": lambda: Workload("ni_c0", _ni(), 20.0, 0.1, 50, 20, 1.0, 20260, 1,
                              decay_nm=15.0, noise=0.05),
    "fe_c1": lambda: Workload("fe_c1", _fe(), 20.0, 0.05, 250, 50, 0.5, 20261, 2,
                              decay_nm=20.0),
    "al2o3_c2": lambda: Workload("al2o3_c2", _al2o3(), 15.0, 0.05, 200, 50, 0.5, 20262, 2,
                                 decay_nm=20.0),
    "fe_c3": lambda: Workload("fe_c3", _fe(), 20.0, 0.05, 150, 60, 2.0, 20263, 2,
                              profile="gamma",
                              energies=tuple(10.0 + 0.5 * i for i in range(21))),
    "ni_c4": lambda: Workload("ni_c4", _ni(), 20.0, 0.1, 100, 20, 1.0, 20264, 1,
                              decay_nm=15.0, noise=0.05),
}


_WIDE_CUTOFFS = {"ni_c0": (240.0, 480.0), "ni_c4": (240.0, 480.0), "fe_c1": (320.0, 640.0),
                 "fe_c3": (320.0, 640.0), "al2o3_c2": (320.0, 640.0)}


# the upper half of corundum's quoted 200-800 strong beams (509-616 measured: matrices past 512)
_WIDER_CUTOFFS = {"al2o3_c2": (640.0, 1280.0)}


def get(name: str) -> Workload:
    if name.endswith("x") and name[:-1] in _WIDER_CUTOFFS:
        wl = _WORKLOADS[name[:-1]]()
        wl.name = name
        wl.weakcutoff, wl.cutoff = _WIDER_CUTOFFS[name[:-1]]
        return wl
    base = name[:-1] if name.endswith("w") and name[:-1] in _WIDE_CUTOFFS else name
    try:
        wl = _WORKLOADS[base]()
    except KeyError:
        raise KeyError(f"unknown workload {name!r}; known: {names()}") from None
    if base != name:
        wl.name = name
        wl.weakcutoff, wl.cutoff = _WIDE_CUTOFFS[base]
    return wl


def names():
    return sorted(list(_WORKLOADS) + [k + "w" for k in _WIDE_CUTOFFS] + [k + "x" for k in _WIDER_CUTOFFS])
