"""A device-resident n-d array: a shape, a dtype and a block of HBM.

Operands that stay on the device between products (the binding's matmul
takes and returns them) are held in a DeviceArray.  Its memory comes from the
library's allocator.  Every product on DeviceArrays is enqueued on the
calling thread's library stream, so results are ready for the next device
call on the same thread at once; `to_numpy` waits for them.  To hand an array
to another thread, call `ferray_amd.sync()` first.
"""
from __future__ import annotations

import numpy as np

from .gemm import DeviceBuffer

__all__ = ["DeviceArray"]

_DTYPES = tuple(np.dtype(t) for t in (np.float64, np.float32, np.complex128, np.complex64,
                                     np.int8, np.uint8, np.int16, np.int32))


class DeviceArray:
    """C-contiguous array in device memory."""

    def __init__(self, shape, dtype, buffer: DeviceBuffer | None = None):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        if self.dtype not in _DTYPES:
            raise TypeError(f"DeviceArray: unsupported dtype {self.dtype}")
        if any(s < 0 for s in self.shape):
            raise ValueError(f"DeviceArray: negative dimension in {self.shape}")
        nbytes = self.size * self.dtype.itemsize
        if buffer is not None and buffer.nbytes < nbytes:
            raise ValueError("DeviceArray: buffer smaller than the array")
        self.buffer = buffer if buffer is not None else DeviceBuffer(nbytes)

    @property
    def ndim(self) -> int:
        return len(self.shape)

    @property
    def size(self) -> int:
        return int(np.prod(self.shape, dtype=np.int64)) if self.shape else 1

    @property
    def nbytes(self) -> int:
        return self.size * self.dtype.itemsize

    @property
    def ptr(self) -> int:
        return self.buffer.ptr

    @classmethod
    def from_numpy(cls, arr) -> "DeviceArray":
        arr = np.ascontiguousarray(arr)
        out = cls(arr.shape, arr.dtype)
        if arr.size:
            out.buffer.upload(arr)
        return out

    def to_numpy(self) -> np.ndarray:
        if self.size == 0:
            return np.zeros(self.shape, self.dtype)
        return self.buffer.download(self.dtype, self.size).reshape(self.shape)

    def reshape(self, *shape) -> "DeviceArray":
        """A view of the same memory with another shape of equal size."""
        if len(shape) == 1 and isinstance(shape[0], (tuple, list)):
            shape = tuple(shape[0])
        view = DeviceArray.__new__(DeviceArray)
        view.shape = tuple(int(s) for s in shape)
        view.dtype = self.dtype
        if view.size != self.size:
            raise ValueError(f"cannot reshape {self.shape} into {view.shape}")
        view.buffer = self.buffer
        return view

    def __repr__(self) -> str:
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"
