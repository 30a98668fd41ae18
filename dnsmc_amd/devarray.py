"""Minimal device array over the library's own allocator (dnsmc_dev_*), so
that GPU-resident training data never needs a second HIP runtime in the
process."""
import ctypes as C

import numpy as np

from ._lib import check, lib


class DeviceArray:
    """Contiguous device buffer with a numpy dtype and shape.  Slicing along
    the first axis returns a view that shares the allocation."""

    def __init__(self, shape, dtype, _ptr=None, _owner=None):
        self.shape = tuple(int(x) for x in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        if _ptr is None:
            p = C.c_void_p()
            check(lib().dnsmc_dev_alloc(self.nbytes, C.byref(p)))
            self.ptr = p.value or 0
            self._owner = None
        else:
            self.ptr = _ptr
            self._owner = _owner

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            check(lib().dnsmc_dev_copy(C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes, 1))
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().dnsmc_dev_copy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, 2))
        return out

    def copy_from(self, src):
        assert src.nbytes == self.nbytes
        check(lib().dnsmc_dev_copy(C.c_void_p(self.ptr), C.c_void_p(src.ptr), self.nbytes, 3))

    @property
    def c(self):
        return C.c_void_p(self.ptr)

    def __len__(self):
        return self.shape[0]

    def __getitem__(self, sl):
        assert isinstance(sl, slice) and sl.step in (None, 1)
        start, stop, _ = sl.indices(self.shape[0])
        row = self.nbytes // max(self.shape[0], 1)
        return DeviceArray((max(stop - start, 0),) + self.shape[1:], self.dtype, _ptr=self.ptr + start * row,
                           _owner=self._owner or self)

    def __del__(self):
        if self._owner is None and getattr(self, "ptr", 0):
            try:
                lib().dnsmc_dev_free(C.c_void_p(self.ptr))
            except Exception:
                pass
            self.ptr = 0


def concat(arrays):
    """Concatenate along the first axis into a new allocation."""
    shape = (sum(a.shape[0] for a in arrays),) + arrays[0].shape[1:]
    out = DeviceArray(shape, arrays[0].dtype)
    off = 0
    for a in arrays:
        if a.nbytes:
            check(lib().dnsmc_dev_copy(C.c_void_p(out.ptr + off), C.c_void_p(a.ptr), a.nbytes, 3))
        off += a.nbytes
    return out
