"""swinmd_amd -- MI355X (gfx950) implementation of SwinMD's nonbonded force path.

The compute path is libswinmd_hip.so (hand-written HIP kernels behind the C ABI
in include/swinmd_hip.h).  Python here is host plumbing: it mirrors the
reference's Force/Integrator classes over that ABI (``swinmd_amd.forces``) and
builds synthetic systems (``swinmd_amd.systems``).
"""
from . import systems  # noqa: F401

__all__ = ["systems", "forces", "build"]
