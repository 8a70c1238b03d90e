"""Low-level HIP operators (same names as the reference's ``r2splat.cuda``)."""

from ._wrapper import *  # noqa: F401,F403
from ._wrapper import __all__  # noqa: F401
