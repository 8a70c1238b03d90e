"""MI355X (gfx950) explicit Gaussian and Beta CT splatting operators.

Drop-in for the public API of r2splat (r2splat/__init__.py:1-20).
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
