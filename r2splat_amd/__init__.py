"""MI355X (gfx950) explicit Gaussian and Beta CT splatting operators.

Drop-in for the public API of r2splat (r2splat/__init__.py:1-20); ``__all__``
is the reference's. The differentiable wrappers ``rasterize`` / ``voxelize``
(torch.autograd.Function over the explicit calls) live in
``r2splat_amd.autograd`` and are not attributes of the package, as in the
reference (whose tests check that ``rasterize`` is absent).
"""

__version__ = "0.1.0"

from .explicit import (
    ParameterGradients,
    rasterize_bwd,
    rasterize_fwd,
    voxelize_bwd,
    voxelize_fwd,
)

__all__ = [
    "__version__",
    "rasterize_fwd",
    "rasterize_bwd",
    "voxelize_fwd",
    "voxelize_bwd",
    "ParameterGradients",
]
