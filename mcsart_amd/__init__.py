"""mcsart_amd -- MI355X-native McSART (motion-compensated SART cone-beam CT reconstruction).

The product is the C-ABI library ``libmcsart_hip.so`` (HIP kernels for gfx950 plus a C++ host
driver).  This package holds its Python host-side mirror (ctypes bindings with the reference's
operator names and argument order), the seeded synthetic-scan generator and the workload
definitions.
"""
from .geometry import Geometry, uniform_angles  # noqa: F401

__all__ = ["Geometry", "uniform_angles"]
