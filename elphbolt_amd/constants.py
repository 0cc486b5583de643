"""Physical constants and unit factors of the vertex path (src/params.f90:27-57)."""
import math

QE = 1.602176634e-19          # C
ME = 9.1093837015e-31         # kg
AMU = 1.66053906660e-27       # kg
HBAR = 1.05457172647e-22      # J/THz
HBAR_EVPS = HBAR / QE
PI = 4.0 * math.atan(1.0)
TWOPI = 2.0 * PI
HARTREE2RADTHZ = 27.2116 * QE / HBAR
HARTREE2EV = 27.2116
RYD2RADTHZ = 0.5 * HARTREE2RADTHZ
RYD2EV = 0.5 * HARTREE2EV
RYD2AMU = 2.0 * ME / AMU
BOHR2NM = 0.052917721092
G2UNITFACTOR = RYD2EV ** 3 * RYD2AMU   # src/wannier.f90:923
