"""goma_amd -- MI355X-native Newton inner loop for Goma's steady Navier-Stokes(+energy) path.

The compute path lives in libgm.so (HIP kernels for gfx950 behind the C ABI of
include/gm_capi.h).  ``goma_amd.capi`` binds it; ``goma_amd.mesh`` and
``goma_amd.problems`` build the synthetic inputs.
"""
__all__ = ["mesh", "problems", "capi"]
