"""marvel_amd: MI355X-native daligner overlap path (MARVEL dalign/ lineage).

The compute lives in the HIP/C++ library `libmarvl_align.so` built in this
directory (see `marvel_amd.build`).  Python here is host plumbing: loading
the library, packing reads, generating synthetic workloads.  Nothing in this
package falls back to CPU arithmetic: if the library or a GPU is missing the
calls raise.
"""

from .lib import lib, LibraryMissing  # noqa: F401

__all__ = ["lib", "LibraryMissing"]
