"""nixis_amd -- MI355X (gfx950) implementation of the Nixis terrain hot path.

Modules named after the reference's (opensimplex, terrain, erosion, util)
provide its functions with the same signatures; they call hand-written HIP
kernels in libnixis_hip.so through a C ABI (include/nixis.h).
"""
__version__ = "0.1.0"


def library_path():
    from ._lib import LIB_PATH
    return LIB_PATH
