"""semhip: MI355X (gfx950) operator-apply-and-solve path of SpectralElements.jl.

A Python mirror of the reference API (Mesh, generateMask, ABu, lapl, mass,
hlmz, gatherScatter, mask, pcg, pcg_, grad, advect, ...) over the C ABI of
libsemhip.so (include/sem_abi.h). See semhip.core for the reference
file:line each function mirrors.
"""
from ._lib import SemError, load as load_library, LIB_PATH  # noqa: F401
from .core import (  # noqa: F401
    ABu, Diagonal, GSMap, MG, Mesh, PCGInfo, SEMOperator, adjoint_dnu, advect, ctx, deform_affine01,
    gordonHall, gordonHallMap, semInterp,
    deform_sinsin, derivMat, empty_field, fixU, gatherScatter, gausslobatto, generateMask, grad,
    gradT, hlmz, identity, interpMat, lapl, mask, mass, ndgrid, norm_inf, pcg, pcg_, pcg_mixed,
    semmesh, synchronize, to_device, to_host, wdot, dot, dot_acc, combine, vardiff_factors, annulus,
)
from .adjoint import (  # noqa: F401
    VarDiffusion, VarDiffusionOperator, diffusion_solver, linsolve, nodal_gradient,
)
from .cd import (  # noqa: F401
    CDGradient, ConvectionDiffusion, Field, TimeStepper, bdfExtK, cd_gradient, cd_objective,
    cd_objective_step,
)
from .diffusion import Diffusion  # noqa: F401
from . import workloads  # noqa: F401
from . import sciml  # noqa: F401
