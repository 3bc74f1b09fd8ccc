"""MI355X-native Cavalry hot path: LBVH build, treelet restructuring, BVH
traversal and the wavefront path-tracing loop as HIP kernels for gfx950,
behind the C ABI of include/cavalry.h."""
from .bvh import BVH, Context  # noqa: F401
