"""frhip -- host side of the MI355X flux-reconstruction solver core.

Mirrors the public API of FluxReconstruction.jl (src/FluxReconstruction.jl:19-51) over the C ABI
of libfrhip.so (include/fr.h).  All numerical work runs in the shared library: operator data on
its native host routines, residuals / stages / dissipation in its gfx950 kernels.
"""
from ._lib import FRError, lib, exported_symbols  # noqa: F401
from .space import (  # noqa: F401
    Line,
    Quad,
    PSpace1D,
    PSpace2D,
    CSpace2D,
    FRPSpace1D,
    FRPSpace2D,
    legendre_point,
    gausslegendre,
    lagrange_point,
    dlagrange,
    standard_lagrange,
    dlegendre,
    dradau,
    dsd,
    dhuynh,
    JacobiP,
    dJacobiP,
    vandermonde_matrix,
    dvandermonde_matrix,
    rs_jacobi,
    rs_jacobi_mesh,
    r_x,
    global_sp,
    positive_limiter,
    modal_filter,
    basis_norm,
    shock_detector,
    L1_error,
    L2_error,
    Linf_error,
)
from .ops import interp_face, poly_derivative, filter_exp  # noqa: F401
from .euler2d import FREuler2D  # noqa: F401
from .fr1d import FR1D, FREulerProblem, FRAdvectionProblem, Ensemble1D  # noqa: F401
from . import ics  # noqa: F401
from .bgk import BGK2D, BGK1D, VSpace1D  # noqa: F401
from . import sensor  # noqa: F401
from . import io  # noqa: F401

__version__ = "0.1.0"
