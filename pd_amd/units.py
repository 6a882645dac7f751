"""PD's physical constants and unit conversions (forcefield.cpp:35-55, maths.cpp:23-25).

PD stores positions in Angstrom, forces in J/m, energies in J per molecule,
masses in kg and per-atom LJ epsilon as sqrt(epsilon[J]) (ffparam.cpp:1216).
The device kernels work in kcal/mol and Angstrom; these helpers convert at the
boundary.
"""
import math

PI = 3.14159265358979323846
TWO_PI = 2.0 * PI
FOUR_PI = 4.0 * PI
PI_OVER_180 = PI / 180.0

KCAL2J = 4184.0
J2KCAL = 1.0 / KCAL2J
NA = 6.0221418e23
KCAL2J_DIV_NA = KCAL2J / NA
KB = 1.38065e-23
AMU = 1.66053878e-27
ANGSTROM = 1e-10
INV_ANGSTROM = 1.0 / ANGSTROM
E_CHARGE = 1.60217645e-19
E0 = 8.854187817e-12
_FOUR_PI_E0 = (PI * 4.0) * E0
ECONV = (J2KCAL * NA * (E_CHARGE * E_CHARGE)) / (_FOUR_PI_E0 * ANGSTROM)  # kcal/mol * A
HALFECONV = ECONV * 0.5
ECONV_JOULE = (E_CHARGE * E_CHARGE) / (_FOUR_PI_E0 * ANGSTROM)  # J * A
HALFECONV_JOULE = ECONV_JOULE * 0.5

# J per molecule -> kcal/mol, and J/m -> kcal/mol/A
J_TO_KCALMOL = J2KCAL * NA
JM_TO_KCALMOLA = J2KCAL * NA * ANGSTROM


def eps_kcal_to_pd(eps_kcal):
    """Per-atom epsilon as PD stores it: sqrt(eps[kcal/mol] * kcal2J / Na) (ffparam.cpp:1216)."""
    return math.sqrt(eps_kcal * KCAL2J / NA)


def kcal_to_j(x):
    """kcal/mol -> J per molecule (force constants, ffparam.cpp:130,233,513)."""
    return x * KCAL2J / NA
