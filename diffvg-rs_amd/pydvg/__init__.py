"""Host side of the MI355X render/backward library (libdvg.so).

`scene` mirrors the crate's Scene / Shape / SceneGrad API, `abi` is the ctypes
view of include/dvg.h, `renderer` drives the C ABI, `scenes` builds the seeded
benchmark workloads.
"""
from . import abi, scene, scenes  # noqa: F401
