"""MI355X (gfx950) energy hot path for GOMC: cell list, real-space LJ +
erfc-Coulomb pair energies, Ewald structure factors and their deltas, the
single-box MC driver and the batched persistent MC kernel, all behind the C
ABI of include/gomc_hip.h (libgomc_hip.so)."""
from .systems import System, SystemDesc  # noqa: F401

__all__ = ["System", "SystemDesc", "Box", "Batch"]


def __getattr__(name):
    # Box and Batch need libgomc_hip.so; loading it is deferred so that the
    # system generators import on machines without the library.
    if name in ("Box", "Batch", "GomcError"):
        from . import api
        return getattr(api, name)
    raise AttributeError(name)
