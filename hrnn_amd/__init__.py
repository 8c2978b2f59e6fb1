"""hrnn_amd -- MI355X-native (gfx950) HRNN hot path.

Layers:
    include/hrnn_dev.h        C-ABI of libhrnn_hip.so (HIP kernels)
    include/hnswlib/hrnn.h    hnswlib::HRNN<float> host API on top of it
    hrnn_amd.dev              Python binding of the C-ABI (ctypes)
"""
from .dev import (PREC_BF16, PREC_F16, PREC_F32, Device, HrnnError, device_count, lib,  # noqa: F401
                  synth_halfnormal_scaled, synth_mixture, synth_sphere_mixture, synth_split_perm)

__all__ = ["Device", "HrnnError", "device_count", "lib", "synth_mixture",
           "synth_halfnormal_scaled", "synth_sphere_mixture", "synth_split_perm", "PREC_F32", "PREC_BF16",
           "PREC_F16"]
