"""deconlib_amd — MI355X-native FFT-operator deconvolution.

Drop-in for the hot path of ``deconlib.deconvolution``: the operator protocol
and composition, ``LinearFFTConvolver`` / ``FFTConvolver``, ``Pad`` / ``Crop`` /
``FractionalAreaDownsample``, ``Gradient1D/2D/3D``, ``Hessian1D/2D/3D``, Richardson–Lucy,
PDHG with identity, gradient and Hessian regularization, NLCG (``nlcg_with_operator``,
``nlcg_solver``), ER-Decon (``erdecon_with_operator``, ``erdecon_solver``), JetNewton
(``jetnewton_with_operator``), and the tiled path (``make_forward_model``,
``plan_tiles`` / ``process_tiles``, ``richardson_lucy_solver``), and ``.psf.h5`` files (``Psf``,
``save_psf``, ``load_psf``, through the system's HDF5 C library).  All arithmetic runs in hand-written HIP kernels for
gfx950 (``deconlib_amd/_lib/libdeconhip.so``, C ABI in ``include/deconhip.h``);
PyTorch only owns device buffers and streams.
"""

from .base import DeconvolutionResult, MLXDeconvolutionResult
from .composition import Compose, LinearOperator, compose
from .core_operators import Crop, FractionalAreaDownsample, Pad
from .linops import (AnisotropicHessian2D, AnisotropicHessian3D, FFTConvolver, Gradient1D, Gradient2D, Gradient3D,
                     Hessian1D, Hessian2D, Hessian3D, LinearFFTConvolver, fast_padded_shape)
from .nlcg import NLCGResult, nlcg_solver, nlcg_with_operator
from .norms import power_iteration_norm_sq
from .erdecon import ERDeconResult, erdecon_solver, erdecon_with_operator
from .jetnewton import JetNewtonResult, jetnewton_with_operator
from .pdhg import (GradientRegularizer, HessianRegularizer, IdentityRegularizer, solve_pdhg_mlx,
                   solve_pdhg_with_operator)
from .rl import RLResult, poisson_i_divergence, richardson_lucy_solver, richardson_lucy_with_operator
from .forward_model import ForwardModel, make_forward_model
from .io import Psf, load_psf, save_psf
from .shapes import compute_padded_shape, compute_visible_shape, get_valid_slices
from .tile_processing import TilePlan, TileSpec, optimal_tile_size, plan_tiles, process_tiles

__all__ = [
    "solve_pdhg_mlx",
    "solve_pdhg_with_operator",
    "IdentityRegularizer",
    "GradientRegularizer",
    "HessianRegularizer",
    "Gradient1D",
    "Gradient2D",
    "Gradient3D",
    "Hessian1D",
    "Hessian3D",
    "Hessian2D",
    "NLCGResult",
    "nlcg_with_operator",
    "nlcg_solver",
    "ERDeconResult",
    "erdecon_with_operator",
    "erdecon_solver",
    "JetNewtonResult",
    "jetnewton_with_operator",
    "AnisotropicHessian2D",
    "AnisotropicHessian3D",
    "MLXDeconvolutionResult",
    "DeconvolutionResult",
    "RLResult",
    "richardson_lucy_with_operator",
    "richardson_lucy_solver",
    "ForwardModel",
    "make_forward_model",
    "Psf",
    "save_psf",
    "load_psf",
    "TileSpec",
    "TilePlan",
    "plan_tiles",
    "process_tiles",
    "optimal_tile_size",
    "compute_padded_shape",
    "compute_visible_shape",
    "get_valid_slices",
    "poisson_i_divergence",
    "FFTConvolver",
    "LinearFFTConvolver",
    "fast_padded_shape",
    "power_iteration_norm_sq",
    "Compose",
    "LinearOperator",
    "compose",
    "Pad",
    "Crop",
    "FractionalAreaDownsample",
]
