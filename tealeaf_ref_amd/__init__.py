"""TeaLeaf for MI355X: hand-written gfx950 HIP kernels behind a C ABI
(include/tealeaf_hip.h), a Fortran host that keeps the reference's tea.in deck
and tea_leaf() driver, and Python bindings (ctypes) over the same library."""

__version__ = "0.1.0"
