"""MI355X-native LenseFlow and lensed Wiener filter (drop-in for CMBLensing.py's
lense_flow / wiener_filter call surface), computing in libcmbl.so (HIP, gfx950).

The directory name contains a dot, so load it with ``load_package()`` from the
repository's ``__graft_entry__`` (registered as module ``cmblensing_amd``).
"""
from .constants import (ARCMIN_PER_DEGREE, Basis, DEFAULT_NSTEPS, FORWARD_LENSE, INVERSE_LENSE,
                        NSIDE_DEFAULT, PIX_WIDTH_DEFAULT, Parametrization, THETA_PIX_DEFAULT)
from .fields import (Field, FlatS0, FlatS02, FlatS2, dot, eb2qu, fourier, fourier_weights_for,
                     map, qu2eb, undo_inner_product)
from .lense_flow import lense_flow, lense_flow_matrices, lense_flow_vjp, lense_flow_wrapper
from ._lib import CmblError, clear_caches

__all__ = [
    "ARCMIN_PER_DEGREE", "Basis", "DEFAULT_NSTEPS", "FORWARD_LENSE", "INVERSE_LENSE",
    "NSIDE_DEFAULT", "PIX_WIDTH_DEFAULT", "Parametrization", "THETA_PIX_DEFAULT",
    "Field", "FlatS0", "FlatS02", "FlatS2", "dot", "eb2qu", "fourier", "fourier_weights_for",
    "map", "qu2eb", "undo_inner_product", "lense_flow", "lense_flow_matrices", "lense_flow_vjp",
    "lense_flow_wrapper", "CmblError",
    "clear_caches",
]
