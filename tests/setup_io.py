"""Writes named arrays to the binary file the C++ backend test reads (tests/cpp/setup_io.hpp):
"S2LSETUP", a record count, then per record: name length (u32), name, type (u8: 0 int32,
1 int64, 2 float64, 3 text), element count (u64), data (little endian)."""
import struct

import numpy as np

_TYPES = {np.dtype(np.int32): 0, np.dtype(np.int64): 1, np.dtype(np.float64): 2}


def write_arrays(path, arrays: dict):
    with open(path, "wb") as f:
        f.write(b"S2LSETUP")
        f.write(struct.pack("<I", len(arrays)))
        for name, v in arrays.items():
            nb = name.encode()
            f.write(struct.pack("<I", len(nb)))
            f.write(nb)
            if isinstance(v, str):
                data = v.encode()
                f.write(struct.pack("<BQ", 3, len(data)))
                f.write(data)
                continue
            a = np.ascontiguousarray(np.atleast_1d(v))
            if a.dtype not in _TYPES:
                a = a.astype(np.float64 if np.issubdtype(a.dtype, np.floating) else np.int64)
            f.write(struct.pack("<BQ", _TYPES[a.dtype], a.size))
            f.write(a.astype(a.dtype.newbyteorder("<")).tobytes())
