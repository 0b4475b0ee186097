"""MI355X-native OPSEC hot path: pair quadrature, KL modes, projection, Fisher.

The compute path is libopsec_hip.so (hand-written HIP kernels for gfx950 behind
the C ABI in include/opsec_hip.h) and the drop-in programs klt, comma and
estimate under opsec_amd/bin.  This package holds the ctypes binding (lib.py),
the .abn/config formats (formats.py) and the synthetic survey configurations
(synth.py).
"""
from . import lib  # noqa: F401
