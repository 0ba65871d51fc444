"""barnacle_amd -- MI355X-native wavefront path tracer for the Barnacle hot path.

The product is the HIP library lib/libbarnacle_hip.so behind the C ABI of
include/barnacle_hip.h; this package holds its sources (csrc/), its build
(build.py), the ctypes binding (native.py), the scene-JSON front end
(scene.py) and the synthetic workload generators (configs.py).
"""
from . import abi, scene, configs  # noqa: F401
