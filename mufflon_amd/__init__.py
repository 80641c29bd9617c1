"""MI355X (gfx950) device of Mufflon's GPU ray-tracing hot path.

mufflon_amd.hip     ctypes binding of the C-ABI in include/mufflon_hip.h
mufflon_amd.scenes  procedural benchmark inputs
mufflon_amd.build   in-tree hipcc build of mufflon_amd/lib/libmufflon_hip.so
"""
