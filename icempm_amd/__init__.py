"""icempm_amd -- the MLS-MPM substep of iceMPM (2D ice indentation) on
MI355X (gfx950): hand-written HIP kernels behind the reference's
GPU_Implementation5 / icy::Model surface.

Native parts (built by ``python -m icempm_amd.build``):
  libicempm_hip.so   kernels + C ABI          (include/icempm.h)
  libicempm_host.so  C++ drop-in surface      (include/icempm_host.h)
"""
from ._lib import (H5_EXPORTS, HIP_EXPORTS, HOST_EXPORTS, MODE_FAST, MODE_ORDERED, IcempmError,
                   SimParams, hip, host)
from .model import (CASES, INDENT_1M6, INDENT_60K, INDENT_TINY, IndentationCase, Model,
                    configure_params, lattice_points, nacc_strengths, soa_views)

__all__ = [
    "H5_EXPORTS", "HIP_EXPORTS", "HOST_EXPORTS", "MODE_FAST", "MODE_ORDERED", "IcempmError", "SimParams",
    "hip", "host", "CASES", "INDENT_1M6", "INDENT_60K", "INDENT_TINY", "IndentationCase",
    "Model", "configure_params", "lattice_points", "nacc_strengths", "soa_views",
]
