"""Physical constants of the SCALE library (scale_const.F90: CONST_GRAV,
CONST_Rdry, CONST_CPdry, CONST_CVdry = CPdry - Rdry, CONST_PRE00).

SCALE is an external dependency of the reference (INSTALL.md) and is not part of
it, so these values are not pinned by anything in the reference tree; every
entry point takes them as inputs, so a host that owns SCALE passes its own."""
from dataclasses import dataclass


@dataclass(frozen=True)
class Constants:
    GRAV: float = 9.80665
    Rdry: float = 287.04
    CPdry: float = 1004.64
    CVdry: float = 1004.64 - 287.04
    PRES00: float = 100000.0


SCALE = Constants()
