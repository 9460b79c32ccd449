"""Host side of the MI355X render/backward library (libdvg.so).

`scene` mirrors the crate's Scene / Shape API and holds the flat gradients,
`grad` the SceneGrad tree and apply_gpu_grads, `abi` is the ctypes view of
include/dvg.h, `renderer` drives the C ABI, `scenes` builds the seeded
benchmark workloads.
"""
from . import abi, grad, scene, scenes  # noqa: F401
