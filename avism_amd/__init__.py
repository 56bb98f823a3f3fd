"""avism_amd -- MI355X-native implementation of AVISM's grid-and-void path.

The compute path is the HIP library ``avism_amd/libavism_hip.so`` (C ABI declared
in ``include/avism.h``).  This package holds the host-side mirror of the
reference's driver for that path: ``voids.dat`` handling, input/output formats,
ctypes bindings and the per-level pipeline.
"""
__version__ = "0.1.0"
