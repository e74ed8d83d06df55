"""MI355X-native (gfx950) implementation of the Morpheus SpMV hot path.

The compute lives in ``libmorpheus_amd.so`` (hand-written HIP kernels behind
the C-ABI of ``include/morpheus_amd.h``); the C++ drop-in headers for the
reference API are under ``morpheus_amd/include``.  This Python package holds
the host containers, the synthetic workload generators, the ctypes binding
and the torch-backed device containers used by tests and the benchmark.

Importing the package does not import torch; ``morpheus_amd.device`` does.
"""
from . import formats  # noqa: F401
from .formats import (COO_FORMAT, CSR_FORMAT, DIA_FORMAT, ELL_FORMAT,  # noqa: F401
                      HDC_FORMAT, HYB_FORMAT, NFORMATS, FORMAT_NAMES)

__version__ = "0.1.0"
