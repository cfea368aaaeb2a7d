"""ORACLE - test infrastructure only.

Reader of the radial-loop dumps written by oracle/ref_rloop_dump.f90 (the
reference's own radial loop, one radialLoopG call: inputs, switches and
outputs).  Records: name (32 bytes), kind (int32: 1 real64, 2 complex128,
3 int32), rank (int32), shape (rank x int32), data in column-major order;
the record named END closes the file.
"""
import numpy as np

_KIND = {1: np.float64, 2: np.complex128, 3: np.int32}


def read_dump(path):
    """dict name -> numpy array (Fortran arrays transposed to C order, so an
    (lm_max, n_r) field reads as [n_r][lm_max]); scalars as 0-d arrays."""
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while True:
        name = data[pos:pos + 32].decode().strip()
        pos += 32
        kind, rank = np.frombuffer(data, dtype="<i4", count=2, offset=pos)
        pos += 8
        shape = tuple(int(s) for s in np.frombuffer(data, dtype="<i4", count=rank, offset=pos))
        pos += 4 * rank
        if name == "END":
            break
        dt = np.dtype(_KIND[int(kind)]).newbyteorder("<")
        n = int(np.prod(shape)) if rank else 1
        arr = np.frombuffer(data, dtype=dt, count=n, offset=pos).reshape(shape, order="F")
        pos += n * dt.itemsize
        arr = np.ascontiguousarray(arr.T)
        if shape == (1,):
            arr = arr.reshape(())
        out[name] = arr
    return out
