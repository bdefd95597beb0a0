"""cgbdsm_amd -- MI355X-native Stage 2 of CGBDSM continuous subgraph matching.

Native parts (built in-tree by cgbdsm_amd.build):
  libcgbdsm.so       gfx950 HIP engine behind the C ABI of include/cgbdsm.h
  libcgbdsm_host.so  C++ host pipeline behind include/cgbdsm_host.h
  bin/cgcsm_*        reference-compatible command line drivers
"""
from . import _native  # noqa: F401

__all__ = ["host", "engine", "build"]
