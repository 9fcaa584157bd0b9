"""MI355X-native parADMM engine: libparadmm.so (C API of paradmm.h + paradmm_gpu.h)."""
from . import paradmm  # noqa: F401
