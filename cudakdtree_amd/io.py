"""The reference's point-file format (cukd::common::loadPoints,
common.h:245-266): a little-endian size_t element count, then the raw
elements (float2/3/4 ... rows of num_dims float32, or any fixed-size
record)."""
import numpy as np


def save_points(path, points):
    """points: (n, D) float32/int32 array or a structured array."""
    a = np.ascontiguousarray(points)
    with open(path, "wb") as f:
        f.write(np.uint64(len(a)).tobytes())
        f.write(a.tobytes())


def load_points(path, num_dims=3, dtype=np.float32):
    """Returns an (n, num_dims) array (dtype may also be a structured dtype,
    then num_dims is ignored and a 1-D record array is returned)."""
    with open(path, "rb") as f:
        head = f.read(8)
        if len(head) != 8:
            raise ValueError("%s: no element count" % path)
        n = int(np.frombuffer(head, np.uint64)[0])
        dt = np.dtype(dtype)
        shape = (n,) if dt.names else (n, num_dims)
        want = int(np.prod(shape)) * dt.itemsize
        body = f.read(want)
        if len(body) != want:
            raise ValueError("%s: file is shorter than its count (%d elements)" % (path, n))
        return np.frombuffer(body, dt).reshape(shape).copy()
