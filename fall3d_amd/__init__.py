"""fall3d_amd: MI355X-native transport core for FALL3D-7.0.

The compute path is libfall3d_amd.so (HIP kernels for gfx950 behind the C ABI
declared in include/fall3d_amd.h); this package holds its ctypes binding
(`lib`), the synthetic scenario generator (`scenario`) and the build recipe
(`build`).
"""
__all__ = ["lib", "scenario", "build"]
