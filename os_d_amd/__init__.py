"""os_d_amd: MI355X-native quantum-jump trajectory engine for the driven,
dissipative Bose-Hubbard dimer (the QJX QJ path of GillianGrayson/os_d).

The engine is the HIP library os_d_amd/_lib/libqjx.so behind the C ABI in
include/qjx.h; os_d_amd.qjx is a thin ctypes mirror of that ABI.
"""
from .qjx import (Engine, Batch, Lindblad, Pool, ModelDesc, EnsembleDesc, QJXError, load_library,  # noqa: F401
                  FP64, FP32, BF16X3, run_config)
