"""Reader for the MAT v4 files of write_to_matlab / bfvm_write_matlab.

The layout (BioFVM_matlab.cpp:515-559): five little-endian 32-bit words
(type, rows, cols, imaginary, name length), the name (no terminating NUL),
then rows x cols FP64 values stored column by column.
"""
import numpy as np


def read_matlab(filename):
    """Returns (name, matrix of shape (rows, cols))."""
    with open(filename, "rb") as f:
        header = np.frombuffer(f.read(20), dtype="<u4")
        mtype, rows, cols, imag, nlen = (int(v) for v in header)
        if mtype != 0 or imag != 0:
            raise ValueError(f"{filename}: not a real little-endian FP64 MAT v4 matrix")
        name = f.read(nlen).decode()
        data = np.frombuffer(f.read(rows * cols * 8), dtype="<f8")
    if data.size != rows * cols:
        raise ValueError(f"{filename}: truncated")
    return name, data.reshape(cols, rows).T
