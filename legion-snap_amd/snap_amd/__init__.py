"""snap_amd: MI355X-native mini-KBA solve (Legion-SNAP hot path).

Host-side mirror of the reference's GPU task entry points over the C ABI in
include/snap_hip.h (libsnap_hip.so, built in-tree for gfx950):

* problems  -- the Problem tables and the workload generators
* deck      -- the reference's input decks (namelist, tables, time stepping)
* solver    -- the Legion-free transport_solve on the GPU (steady, time
               dependent, MMS), gradient of a detector response
* batch     -- many uneven fixed-source problems per launch
* tasks     -- one wrapper per reference GPU task body (run_* entry points)
* kba       -- transport_solve with the decks' chunk decomposition over tasks
* run_deck  -- `python -m snap_amd.run_deck <deck>`, the reference executable's
               report
* autograd  -- solve + gradient as a torch.autograd.Function
"""
from . import problems  # noqa: F401

__all__ = ["problems", "deck", "solver", "batch", "tasks", "kba", "run_deck", "autograd"]
__version__ = "0.1.0"
