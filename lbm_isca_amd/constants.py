"""Physical constants and the T42 L20 resolution of the reference
(``pylbm_isca/constants.py:38-109``), kept identical so inputs built for the
reference drop in unchanged."""

import numpy as np

ER = 6.371e6
GRAV = 9.80
RDGAS = 287.04
RVGAS = 461.50
KAPPA = 2.0 / 7.0
CP = RDGAS / KAPPA
OMEGA = 7.2921150e-5
RAIR = RDGAS
EPSV = RDGAS / RVGAS
EPSVT = RVGAS / RDGAS - 1.0
DAYLEN_SEC = 86400.0
# history output: Z gets ps' [hPa] * 100 / (rho0 g) (output.py:11-15)
RHO00 = 1.225
GRAV_GT2GR = 9.80
S2D = 86400.0

IMAX = 128
JMAX = 64
KMAX = 20
NMAX = 42
NSPEC = (NMAX + 1) * (NMAX + 2) // 2

# L20 half levels, index 0 = surface (pylbm_isca/constants.py:92-97)
SIGM = np.array([1.000, 0.990, 0.970, 0.930, 0.870, 0.790, 0.700, 0.600, 0.500, 0.410,
                 0.330, 0.260, 0.200, 0.150, 0.100, 0.070, 0.050, 0.040, 0.030, 0.020,
                 0.000], dtype=np.float64)
PLEV17 = np.array([1000., 925., 850., 700., 600., 500., 400., 300., 250., 200., 150., 100.,
                   70., 50., 30., 20., 10.], dtype=np.float64)
