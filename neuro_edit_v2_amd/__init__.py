"""MI355X (gfx950) simulation backend for neuro_edit_v2's `sim` crate.

The product is the C-ABI library ``_build/libneurohip.so`` (declared in
``include/neuro_hip.h``); this package is the thin host binding used by the
tests and the benchmark:

* :mod:`neuro_edit_v2_amd.sim`   -- ``Sim`` / ``SimConfig``: the reference
  ``Backend`` trait (sim/src/lib.rs:119-184) over the HIP path;
* :mod:`neuro_edit_v2_amd.net`   -- network generators (net-format builders
  and the benchmark workloads) producing the flat simulator layout;
* :mod:`neuro_edit_v2_amd.batch` -- many networks packed into one launch;
* :mod:`neuro_edit_v2_amd.serve` -- continuous-batching request loop.

There is no CPU fallback: when the library or a GPU is missing every entry
point raises.
"""
from .sim import Sim, SimConfig, available  # noqa: F401
