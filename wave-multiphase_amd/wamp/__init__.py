"""wamp -- MI355X-native five-equation stiffened-gas solver step (HIP, gfx950)."""
from .config import (EOS, Grid, Scheme, CaseConfig, SCHEME_SC, SCHEME_M10, SCHEME_MP5,
                     SCHEMES, PREC_FP64, PREC_MIXED)

__all__ = ["EOS", "Grid", "Scheme", "CaseConfig", "SCHEME_SC", "SCHEME_M10", "SCHEME_MP5",
           "SCHEMES", "PREC_FP64", "PREC_MIXED"]
