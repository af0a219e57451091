"""spparks_amd -- SPPARKS' strict-mode Metropolis colour sweep on MI355X.

The compute path is libspk_hip.so (hand-written HIP kernels behind the C ABI in
include/spk_hip.h); this package holds its ctypes binding, the build script
and the Python side of the batched / serving drivers.
"""
from .hip import (APPS, LATTICES, HipError, Lattice, checksum, load_library,  # noqa: F401
                  master_draws, metropolis_table, ncolors,
                  strict_init_params)

__all__ = ["APPS", "LATTICES", "HipError", "Lattice", "checksum", "load_library",
           "master_draws", "metropolis_table", "ncolors", "strict_init_params"]
