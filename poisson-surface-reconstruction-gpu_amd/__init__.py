"""MI355X (gfx950) Poisson surface reconstruction: HIP kernels behind a C ABI.

The directory name is not a Python identifier; load it with ``load_package``
from ``__graft_entry__`` (it registers the package as ``psr_gpu_amd``).
"""
from . import binding, serve, synth  # noqa: F401
from .binding import PoissonReconstruction  # noqa: F401
