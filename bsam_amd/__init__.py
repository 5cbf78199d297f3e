"""MI355X FAS multigrid engine for block-structured AMR Cahn-Hilliard solves.

The engine is ``libbsamgpu.so`` (C ABI in include/bsam_gpu.h, HIP kernels
for gfx950); this package holds its Python binding (:mod:`.engine`) and the
host-side hierarchy geometry (:mod:`.layout`).
"""

from .layout import Layout, Patch, auto_minlevel  # noqa: F401
