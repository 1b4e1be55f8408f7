"""draugr_hip — MI355X (gfx950) backend of Draugr.jl's AMG solve / resetup path.

Host-side mirror of the reference API over the C-ABI in include/draugr_hip.h
(the library is lib/libdraugr_hip.so, built from csrc/).  Importing this
package loads the library; there is no CPU fallback.
"""
from ._lib import DraugrError, LIB_PATH, last_error, lib
from .config import (AbsoluteStrength, AggregationCoarsening, AggressiveCoarsening, AMGConfig,
                     ChebyshevSmootherType, ColoredGaussSeidelType, DirectInterpolation, ExtendedIInterpolation,
                     HMISCoarsening, ILU0SmootherType, JacobiSmootherType, L1ColoredGaussSeidelType,
                     L1JacobiSmootherType, L1SerialGaussSeidelType, PMISCoarsening, RSCoarsening,
                     SerialGaussSeidelType, SerialILU0SmootherType, SignedStrength, SmoothedAggregationCoarsening,
                     SPAI0SmootherType, SPAI1SmootherType, StandardInterpolation, hypre_default_config)
from .core import (AMGHierarchy, DraugrPreconditioner, amg_cycle_, amg_resetup_, amg_setup, amg_solve_, as_csr,
                   aslinearoperator,
                   device_count, ldiv_, pcg_solve_, preconditioner_apply_, preconditioner_update_, upload)
from . import adjoint, batch, problems
from .batch import AMGBatch, amg_batch_setup

lib()  # load now: fail loudly if the HIP library is missing
