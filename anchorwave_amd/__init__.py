"""MI355X (gfx950) implementation of AnchorWave's KSW2 two-piece-affine
aligner ``ksw_extd2``: HIP kernels behind the reference's C ABI
(``include/ksw2_hip.h``), a ctypes binding (``ksw2``) and seeded synthetic
workloads (``workloads``)."""
from . import workloads  # noqa: F401

__all__ = ["ksw2", "workloads"]
