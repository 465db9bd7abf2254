"""gfx950 ISPH step for the porous beach-slope solver.

The compute path is libisph.so (HIP, built for gfx950); this package holds
its ctypes binding (:mod:`.capi`), the parameter struct (:mod:`.params`),
the synthetic workload generators (:mod:`.configs`) and the adjoint wrapper
(:mod:`.adjoint`).
"""
from .params import IsphParams, make_params  # noqa: F401
