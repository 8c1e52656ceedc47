"""MI355X-native IMEXLBM hot path: D3Q27 BGK collide + pull stream + periodic
ghost fill + conserved reductions, as hand-written HIP behind a C ABI
(include/imexlbm.h).  ``capi`` binds that ABI; ``build`` compiles it."""
from . import capi  # noqa: F401

__all__ = ["capi"]
