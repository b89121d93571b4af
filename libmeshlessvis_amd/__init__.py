"""MI355X-native Fourier-projection rendering of meshless datasets.

The product is the C-ABI library ``libmeshless_vis.so`` (HIP kernels for
gfx950 behind the API of the reference's include/meshless.h and
include/meshless_vis.h).  This package holds its sources (``csrc/``), the
in-tree build (``build.py``) and a thin ctypes host layer that mirrors the C
API one-to-one (``abi``) plus a numpy convenience layer (``host``).

Importing the package does not load the library; ``abi.lib()`` does, and
raises if it was not built.  Nothing here falls back to a CPU path.
"""
from .abi import (  # noqa: F401
    BASIS,
    FORM,
    PHI_MODE,
    VisError,
    lib,
    library_path,
)
from .host import Config, Dataset, render  # noqa: F401
