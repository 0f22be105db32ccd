"""rayspace3d_amd -- MI355X-native ray-cast spatial join (RaySpace3D's query path).

The compute lives in librs3d.so (HIP kernels for gfx950 behind the C ABI of
include/rs3d.h). This package holds its Python binding, seeded input generators and the
.pre geometry format."""
from . import _lib, gen  # noqa: F401

__all__ = ["_lib", "gen"]
