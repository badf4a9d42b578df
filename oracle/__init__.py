"""CPU FP64 oracle for the PyLBM-Isca time-stepping core (TEST INFRASTRUCTURE).

This package is a clean-room numpy restatement of the reference's algorithm
for the hot path (spectral transforms, Simmons-Burridge vertical operators,
linearised tendencies, semi-implicit leapfrog with Robert-Asselin filter,
member batching).  It exists only to *check* the HIP product:

* only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
  ``bench.py`` may import it, and only as the checker or the timed CPU
  baseline -- never as the thing measured or shipped;
* the product package ``lbm_isca_amd`` never imports it.

Each function cites the reference file:line it restates (paths relative to
the reference checkout, e.g. ``pylbm_isca/spectral.py:93-154``).

Pinning (what checks the restatement itself, see ``tests/test_oracle_*.py``):

* ``sht``      -- quadrature exactness on monomials, discrete
  orthonormality, closed-form P_n^m / H_n^m at n <= 7, P_n^m at every n of
  T42 and T127 against scipy's spherical harmonics and H_n^m against the
  textbook derivative identity, brute-force DFT of a few rows
  (``tests/test_oracle_sht_pins.py``); the reference's analytic transform
  tests (round trip, Laplacian eigen, gradients, solid body, div(grad)=lap)
  and reference-generated vectors (``tests/golden``).
* ``vgrid``    -- closed forms and the hydrostatic recurrence of the
  reference tests, plus reference-generated operator values.
* ``dyn``      -- exact zero on a resting isothermal state, tangent-linear
  Taylor slope, and reference-generated tendencies on a baroclinic jet.
* ``integ``    -- semi-implicit residual/conditioning, reference-generated
  step output with active damping and vertical diffusion.
* ``model``    -- the reference-held Gill golden (gill_isca.json) and
  notebook maxima, reference-generated multi-step samples, ensemble
  fold semantics (identical members == single run); linear mode by
  superposition and first-order agreement with the strictly linear model.
* ``gauss_grid(refine=True)`` -- extended-precision weights (not the
  reference's ``leggauss``), checked against 40-digit Newton (mpmath).
"""

from .sht import SHT, gauss_grid, legendre_tables  # noqa: F401
from .vgrid import VerticalOps  # noqa: F401
from .dyn import BasicTerms, tendencies  # noqa: F401
from .integ import Damping, Stepper  # noqa: F401
from .model import OracleConfig, OracleModel, GridState  # noqa: F401
