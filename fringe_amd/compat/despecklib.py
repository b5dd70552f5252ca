"""Top-level alias so drivers written against the reference bindings
(`import despecklib`) load this library: put fringe_amd/compat on PYTHONPATH."""
from fringe_amd.despecklib import *  # noqa: F401,F403
