"""MI355X (gfx950) implementation of the reference software renderer's hot
path: GPU LBVH build, BVH4 closest-hit traversal, wavefront diffuse path
tracing with a path-replay adjoint.

The compute path is libmtrt.so (HIP, C ABI in include/rt.h); the scene /
camera / config layer is libmtrt_host.so (C++, include/rt_host.h).  Both are
built in-tree by master_thesis_amd.build; importing the bindings raises when
they are missing — there is no fallback path.
"""
from ._lib import RTError  # noqa: F401
from .scenes import Scene, cornell_box, triangle_soup, microbench_rays  # noqa: F401

__all__ = ["RTError", "Scene", "cornell_box", "triangle_soup", "microbench_rays"]
