"""MI355X-native vertex-centric iteration engine (host-side Python mirror).

The engine itself is the header-only C++/HIP template library under
``kites_amd/csrc/kites`` (drop-in for the reference's ``kites/`` headers);
this package binds its C ABI (``include/kites_amd.h``) for tests, the
benchmark and the scheduler trainer.
"""
from .graph import (DeviceGraph, HostGraph, autotune_vpg, device_count, mlp_nparam,  # noqa: F401
                    run_opts)
from ._lib import ALGO, INF, SCHED, VARIANTS, KitesError  # noqa: F401
