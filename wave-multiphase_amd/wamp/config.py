"""Run-time configuration of one solver case: grid, EOS, scheme, boundaries.

The reference hard-codes these as module constants (sainath_SC_it_works.py
SC:25-70, Sainath_M10.py M10:24-67); here every one is a run-time parameter
so the same step serves the SC variant, the M10 variant and the MP5/central6
combination assembled from helpers the reference defines (SC:353-412).
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace

G_DEFAULT = 5  # ghost layers, SC:26

# Scheme switch values (shared with include/wamp.h).
BASE_MUSCL, BASE_MP5 = 0, 1
SHEAR_NONE, SHEAR_CENTRAL4, SHEAR_CENTRAL6 = 0, 1, 2
VF_GATED, VF_ALWAYS = 0, 1
FACE_PP_SIMPLE, FACE_PP_WONG = 0, 1
BC_ZERO_GRADIENT, BC_SLIP_WALL = 0, 1

PREC_FP64, PREC_MIXED = 0, 1


@dataclass(frozen=True)
class EOS:
    """Two stiffened-gas phases: (gamma, p_inf) for phase 1 and phase 2 (SC:40-43)."""

    gamma1: float = 6.12
    pinf1: float = 3.43e8
    gamma2: float = 1.4
    pinf2: float = 0.0

    def as_tuple(self):
        return (self.gamma1, self.pinf1, self.gamma2, self.pinf2)


@dataclass(frozen=True)
class Scheme:
    """Reconstruction / limiter switches.

    base      MUSCL kappa=1/3 (SC:1061-1073) or MP5 (SC:353-405)
    shear     none (M10), central4 (SC:1074-1082) or central6 (SC:407-412),
              applied to the shear wave where the Ducros sensor is < 0.01
    vf_thinc  THINC on the volume-fraction wave gated by the entropy indicator
              (SC:1085-1092) or always (M10:971-979)
    face_pp   simple first-order fallback (SC:1102-1115) or Wong staged face
              limiter with pressure floor (M10:1007-1025)
    repair    USE_EMERGENCY_REPAIR (SC:37)
    """

    base: int = BASE_MUSCL
    shear: int = SHEAR_CENTRAL4
    vf_thinc: int = VF_GATED
    face_pp: int = FACE_PP_SIMPLE
    repair: bool = True
    cfl: float = 0.4  # SC:34


SCHEME_SC = Scheme(BASE_MUSCL, SHEAR_CENTRAL4, VF_GATED, FACE_PP_SIMPLE, True, 0.4)
SCHEME_M10 = Scheme(BASE_MUSCL, SHEAR_NONE, VF_ALWAYS, FACE_PP_WONG, True, 0.4)
SCHEME_MP5 = Scheme(BASE_MP5, SHEAR_CENTRAL6, VF_GATED, FACE_PP_SIMPLE, True, 0.4)

SCHEMES = {"sc": SCHEME_SC, "m10": SCHEME_M10, "mp5": SCHEME_MP5}


@dataclass(frozen=True)
class Grid:
    """Uniform cell-centred grid with G ghost layers; x_i = xmin + (i-G+1/2)dx (SC:1832)."""

    nx: int
    ny: int
    xmin: float = 0.0
    xmax: float = 1.0
    ymin: float = 0.0
    ymax: float = 1.0
    g: int = G_DEFAULT
    # boundary types: x-lo, x-hi, y-lo, y-hi (SC:1616-1682: zero-gradient x, slip-wall y)
    bc: tuple = (BC_ZERO_GRADIENT, BC_ZERO_GRADIENT, BC_SLIP_WALL, BC_SLIP_WALL)

    @property
    def nxg(self) -> int:
        return self.nx + 2 * self.g

    @property
    def nyg(self) -> int:
        return self.ny + 2 * self.g

    @property
    def dx(self) -> float:
        return (self.xmax - self.xmin) / self.nx

    @property
    def dy(self) -> float:
        return (self.ymax - self.ymin) / self.ny

    def x_centres(self):
        import numpy as np

        return np.array([self.xmin + (i - self.g + 0.5) * self.dx for i in range(self.nxg)],
                        dtype=np.float64)

    def y_centres(self):
        import numpy as np

        return np.array([self.ymin + (j - self.g + 0.5) * self.dy for j in range(self.nyg)],
                        dtype=np.float64)

    def interior(self):
        return (slice(self.g, self.g + self.nx), slice(self.g, self.g + self.ny))


@dataclass
class CaseConfig:
    """Everything one solve needs besides its initial state."""

    grid: Grid
    eos: EOS = field(default_factory=EOS)
    scheme: Scheme = SCHEME_SC
    precision: int = PREC_FP64
    name: str = "case"

    def with_precision(self, precision: int) -> "CaseConfig":
        return replace(self, precision=precision)
