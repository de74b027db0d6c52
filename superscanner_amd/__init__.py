"""MI355X-native voxel-hashed TSDF engine (allocate, integrate/deintegrate,
raycast) with the call surface of SuperScanner's GPUVoxelGrid.

The compute path is libtsdf_hip.so (hand-written HIP for gfx950) behind the C
ABI of include/tsdf.h; this package holds its ctypes binding, a host-side
mirror of the reference class, and the synthetic-frame generator.
"""
