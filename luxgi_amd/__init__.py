"""luxgi_amd -- MI355X (gfx950) implementation of LuxGI's global-illumination hot path.

The product is libluxgi.so (hand-written HIP kernels behind the C ABI in
include/luxgi.h). This package holds its build recipe (build.py), the ctypes
mirror of the ABI structs (abi.py), the loader (_lib.py, fails loudly when the
library is missing), host mirrors of the reference's ECS systems (ddgi.py,
svgf.py) and synthetic scene builders (scene.py).
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
