"""snap_amd: MI355X-native steady-state mini-KBA solve (Legion-SNAP hot path).

Host-side mirror of the reference's GPU task entry points over the C ABI in
include/snap_hip.h (libsnap_hip.so, built in-tree for gfx950).
"""
from . import problems  # noqa: F401

__all__ = ["problems", "solver", "tasks"]
__version__ = "0.1.0"
