"""lbm_isca_amd -- the PyLBM-Isca time-stepping core on the AMD MI355X (gfx950).

A ``device="hip"`` backend with the reference package's API
(``pylbm_isca/__init__.py:21-32``): ``Spectral``, ``VerticalGrid``, ``LBM``,
``Config``, ``State``, ``make_ensemble_model``.  The hot path runs in
hand-written HIP kernels (``lbm_isca_amd/csrc``) reached through the C ABI
of ``include/lbm_hip.h``; PyTorch-ROCm only holds device memory and streams.
History writers (``output``), forcing readers (``forcing``), spectral
regridding (``regrid``) and the basic-state pipeline (``basicstate``) keep
the reference's module names.
"""

from .backend import asnumpy, device_name, gpu_available, resolve, synchronize
from .constants import IMAX, JMAX, KMAX, NMAX
from .ensemble import BatchedSpectral, EnsembleForcing, build_heating_sweep, make_ensemble_model
from .forcing import ConstForcing, MJOForcing, SteadyForcing, autodetect, read_fortran_records
from .model import LBM, Config
from .output import GradsWriter, MultiWriter, NetCDFWriter
from .spectral import Spectral
from .state import State
from .vertical import VerticalGrid

__all__ = ["Spectral", "VerticalGrid", "LBM", "Config", "State", "BatchedSpectral",
           "EnsembleForcing", "ConstForcing", "SteadyForcing", "MJOForcing", "autodetect",
           "read_fortran_records", "GradsWriter", "NetCDFWriter", "MultiWriter",
           "build_heating_sweep",
           "make_ensemble_model", "IMAX", "JMAX", "KMAX", "NMAX", "gpu_available",
           "device_name", "asnumpy", "resolve", "synchronize"]
__version__ = "0.1.0"
