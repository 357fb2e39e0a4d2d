"""jax_fli_amd: an MI355X-native particle-mesh engine with the public API of jax-fli's PM path.

Imported as ``jax_fli_amd`` (the directory is ``jax-fli_amd/``; see ``__graft_entry__.load_package``).
The compute path is libpmx.so (hand-written HIP for gfx950, C ABI in include/pmx.h); PyTorch only
provides device memory and streams.
"""

from . import _lib, cosmology, engine

__version__ = "0.1.0"
