"""MI355X-native DQMC sweep engine for the O(3) spin-fermion model.

The compute path is libdqmc_hip.so (hand-written HIP for gfx950, C ABI in
include/dqmc_hip.h); this package is its host-side mirror of the reference's
Julia interface (DQMC, init!, propagate, local_updates, ...).
"""
from ._capi import DQMCError, LIB_PATH, load  # noqa: F401
from .model import Params, build_operators, neighbor_table, params_from_dict, params_from_xml  # noqa: F401


def __getattr__(name):
    # DQMC/Batch load the HIP library on first use, so `import dqmc_amd` works on
    # machines without a GPU (setup, parameter parsing, ABI inspection).
    if name in ("DQMC", "Batch", "UniformStream"):
        from . import dqmc
        return getattr(dqmc, name)
    raise AttributeError(name)
