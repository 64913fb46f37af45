"""bajo_amd: MI355X-native (gfx950) BVH and ray-tracing hot path with bajo's API.

Submodules:
  sort    -- bajo/sort/gpu   (device_radix_sort_*, onesweep_radix_sort_*)
  bvh     -- bajo/bvh/gpu    (build_triangle_bvh, GpuTriangleBvh.launch_*)
  scenes  -- seeded synthetic scenes for the benchmark workloads

Everything runs in hand-written HIP kernels in libbajo_hip.so; importing a
submodule that needs the library raises if it has not been built.
"""

from ._lib import BajoError, LIB_PATH  # noqa: F401

__all__ = ["BajoError", "LIB_PATH"]
