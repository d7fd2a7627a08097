"""forceatlas2_rc_amd — ForceAtlas2 for AMD Instinct MI355X (gfx950).

The GPU hot path of forceatlas2r (repulsion offload and all-GPU layout)
rebuilt as HIP kernels behind the reference's C ABI (include/fa2.h), plus
batched layouts of many graphs and a layout server (include/fa2_hip.h).
"""
from .api import (Batch, GPUError, Server, gpu_devices, layout_forceatlas2_gpu,
                  layout_forceatlas2_large, prepare_large, prepare_weights,
                  serve_footprint)

__all__ = ["Batch", "GPUError", "Server", "gpu_devices", "layout_forceatlas2_gpu",
           "layout_forceatlas2_large", "prepare_large", "prepare_weights",
           "serve_footprint"]
