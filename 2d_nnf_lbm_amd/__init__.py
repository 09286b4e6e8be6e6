"""MI355X-native D2Q9 MRT / smoothed-profile time step of taha-re/2D_NNF_LBM.

The product is the HIP library liblbm_amd.so (C ABI in include/lbm_amd.h);
this package holds its ctypes binding, the host-side mirror of the reference
subroutine interface (solver.D2Q9) and the run configurations.  Import the
package with importlib.import_module("2d_nnf_lbm_amd").
"""
from . import _abi  # noqa: F401
from .solver import Batch, D2Q9, LBMError, make_params  # noqa: F401

__all__ = ["Batch", "D2Q9", "LBMError", "make_params"]
