"""simpeg_dc_torch_amd — MI355X-native forward and adjoint for DC resistivity.

Drop-in for the hot path of ``dctorch`` (eacuellarq/SimPEG_DC_Torch 0.2.0):
the same engine and solver interfaces (dctorch/__init__.py:12-21), backed on
the GPU by libdcx.so (hand-written gfx950 HIP kernels behind a C-ABI, loaded
with ctypes) and on the CPU by SciPy SuperLU as in the reference.
"""

__version__ = "0.1.0"

from .solver import (CuDSSBatchSolver, CuDSSSolver, HipBatchSolver,  # noqa: F401
                     HipSolver, SciPyBatchSolver, SciPySolver, make_batch_solver,
                     make_solver, sparse_solve, sparse_solve_batch,
                     sparse_solve_batch_prefactored)
from .simulation2d import TorchDC2D  # noqa: F401
from .linearization2d import Linearization2D  # noqa: F401
from .simulation3d import TorchDC3D  # noqa: F401
from .ip import TorchCR2D, TorchDCIP2D, TorchIP2D, TorchIP3D  # noqa: F401
from .mesh import TensorMesh  # noqa: F401
from . import survey  # noqa: F401
from . import gauss_newton  # noqa: F401
from .lbfgs import lbfgs_minimize  # noqa: F401
