"""specfab_amd: MI355X-native batched fabric evolution and homogenization.

The compute path is libspecfab_gpu.so (HIP, gfx950; C-ABI in
include/specfab_gpu.h), loaded by `specfab_amd._lib`.  `specfab_amd.specfabpy`
mirrors the reference's `specfabpy.specfab` module API on top of it, and
`specfab_amd.batch` is the batched (device-resident) interface.
"""
__all__ = ["_lib", "specfabpy", "batch"]
