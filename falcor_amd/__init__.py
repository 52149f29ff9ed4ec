"""fhp: MI355X-native path-tracing core behind Falcor's render-pass API.

The compute path is libfhp.so (hand-written HIP kernels for gfx950 behind a C ABI, see
include/fhp.h); this package is the Python render-graph surface over it.
"""
from . import scene  # noqa: F401

__all__ = ["scene", "render_graph", "autograd", "image_io", "image_compare", "pyscene"]
__version__ = "0.1.0"
