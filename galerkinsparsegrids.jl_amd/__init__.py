"""MI355X (gfx950) drop-in for GalerkinSparseGrids.jl's hot path.

Mirror of the reference's Julia API for the operator path: the D-dimensional DG
derivative/flux operator (multidim_derivative.jl) applied to the vhier coefficient
vector inside explicit time stepping (pdes.jl).  Every call goes through the C ABI of
libgsg.so (include/gsg.h); there is no CPU fallback.

The directory name contains a dot, so load it with `load_package()` from
__graft_entry__ (it registers the package as `galerkinsparsegrids_jl_amd`).
"""
from ._lib import GsgError, lib
from .ics import (coeffs_DG, cos_coeffs, eval_v, fourier_coeffs, get_one_modal, get_xi_modal, sin_coeffs, tensor_construct,
                  tensor_construct_device, vcoeffs_1d, vcoeffs_DG, vcoeffs_DG_device)
from .layout import D2V, D2Vref, V2D, V2Dref, V2Dref_arrays, get_size
from .operators import (D_matrix, D_matrix_csc, Operator, csc_matrix, dg_basis, grad_matrix, grad_sum_matrix,
                        laplacian_matrix, periodic_DLF_matrix, transport_matrix, wave_matrix)
from .nodal import (eval_points_1D, get_r2_modal, get_r2_point, get_xi2_modal, get_xi_point, load_vlasov_output,
                    make_modal2point_matrices, make_point2modal_matrices, save_vlasov_output, transform, transform_1D,
                    transform_csc, vlasov_evolve, vlasov_operator)
from .serving import Server
from .solvers import (Batch, energy_func, energy_func_1D, norm_squared, transport_evolve, wave_evolve,
                      wave_evolve_1D)

__all__ = [
    "GsgError", "lib", "get_size", "V2Dref", "D2Vref", "V2D", "D2V", "V2Dref_arrays",
    "D_matrix", "grad_matrix", "laplacian_matrix", "grad_sum_matrix", "wave_matrix", "transport_matrix",
    "csc_matrix", "Operator", "dg_basis", "periodic_DLF_matrix", "D_matrix_csc",
    "coeffs_DG", "vcoeffs_DG", "vcoeffs_DG_device", "vcoeffs_1d", "tensor_construct", "tensor_construct_device", "cos_coeffs", "sin_coeffs", "fourier_coeffs",
    "get_one_modal", "get_xi_modal", "eval_v",
    "wave_evolve", "wave_evolve_1D", "transport_evolve", "energy_func", "energy_func_1D", "norm_squared",
    "Batch", "Server",
    "transform_1D", "eval_points_1D", "transform", "make_modal2point_matrices", "make_point2modal_matrices",
    "get_xi_point", "get_xi2_modal", "get_r2_modal", "get_r2_point", "vlasov_operator", "vlasov_evolve",
    "transform_csc", "save_vlasov_output", "load_vlasov_output",
]
