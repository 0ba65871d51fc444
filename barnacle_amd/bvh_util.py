"""Numpy view of the 32-byte BVH node (Util/BVH.fs:52-73)."""
import ctypes as C

import numpy as np

NODE_DTYPE = np.dtype([("pmin", np.float32, 3), ("pmax", np.float32, 3), ("right_or_offset", np.int32),
                       ("is_leaf", np.uint8), ("split_axis", np.int8), ("count", np.int8), ("mask", np.uint8)])


def nodes_to_numpy(nodes, n):
    raw = C.string_at(C.addressof(nodes), NODE_DTYPE.itemsize * n)
    return np.frombuffer(raw, NODE_DTYPE).copy()


def numpy_to_nodes(arr):
    """The inverse of nodes_to_numpy: a ctypes BVHNode array of the same bytes."""
    from .abi import BVHNode
    a = np.ascontiguousarray(arr, NODE_DTYPE)
    return (BVHNode * max(1, len(a))).from_buffer_copy(a.tobytes() if len(a) else bytes(32))
