"""distbvh_amd -- MI355X (gfx950) collision pipeline with distBVH's API.

The device work lives in libdistbvh.so (HIP, built by distbvh_amd/build.py);
`collision` mirrors the reference's collision_world / collision_object over
its C ABI (include/distbvh.h); `workloads` generates the seeded synthetic
inputs.
"""
from . import capi  # noqa: F401
from .collision import CollisionWorld, CollisionObject, CLUSTERING, GEOM_AXIS, ML_GEOM_AXIS, GIVEN  # noqa: F401
