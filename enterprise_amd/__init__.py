"""enterprise_amd: MI355X-native (gfx950, wave64 HIP) direction-optimizing BFS,
a drop-in for iHeartGraph/Enterprise's single-source GPU traversal, with a
batched multi-source form and a continuous-batching serving loop."""
from ._lib import EntError, lib  # noqa: F401
from .graph import HostGraph  # noqa: F401
from .bfs import DeviceGraph, default_config, bfs_gpu_coalescing_mem  # noqa: F401
