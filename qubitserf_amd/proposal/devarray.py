"""Minimal device arrays for the scorer, on the library's own HIP runtime (qsf_dev_*):
float32 or bfloat16 (stored as uint16 bit patterns), row-major, contiguous. Blocks come
from a size-class pool, since a forward pass makes a few dozen short-lived tensors."""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _native

F32, BF16 = "f32", "bf16"
U8, I32 = "u8", "i32"          # raw inputs of the device packing (bytes, int32)
_ITEMSIZE = {F32: 4, BF16: 2, U8: 1, I32: 4}


def _lib():
    L = _native.load()
    if not getattr(L, "_dev_bound", False):
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        L.qsf_dev_alloc.restype = vp
        L.qsf_dev_alloc.argtypes = [sz]
        L.qsf_dev_free.restype = None
        L.qsf_dev_free.argtypes = [vp]
        L.qsf_dev_copy.restype = ctypes.c_int
        L.qsf_dev_copy.argtypes = [vp, vp, sz, ctypes.c_int]
        L.qsf_dev_zero.restype = ctypes.c_int
        L.qsf_dev_zero.argtypes = [vp, sz]
        L.qsf_dev_sync.restype = ctypes.c_int
        L.qsf_dev_sync.argtypes = []
        L._dev_bound = True
    return L


def bf16_bits(x) -> np.ndarray:
    """float32 values -> bfloat16 bit patterns (round to nearest even)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def bf16_to_f32(b) -> np.ndarray:
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


_POOL: dict = {}          # rounded byte size -> free device blocks (reused, never freed)


def _take(nbytes):
    size = 256
    while size < nbytes:
        size <<= 1
    free = _POOL.setdefault(size, [])
    if free:
        return free.pop(), size
    p = _lib().qsf_dev_alloc(size)
    if not p:
        raise MemoryError(f"device allocation of {size} bytes failed")
    return p, size


class DevArray:
    def __init__(self, shape, dtype=F32, zero=True):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = dtype
        self.numel = int(np.prod(self.shape)) if self.shape else 1
        self.itemsize = _ITEMSIZE[dtype]
        self.nbytes = self.numel * self.itemsize
        self.L = _lib()
        self.ptr, self._block = _take(max(1, self.nbytes))
        if zero and self.nbytes:
            self._ok(self.L.qsf_dev_zero(self.ptr, self.nbytes), "qsf_dev_zero")

    @staticmethod
    def _ok(rc, what):
        if rc:
            raise RuntimeError(f"{what} failed (rc={rc})")

    @classmethod
    def from_numpy(cls, a, dtype=F32):
        a = np.asarray(a, np.float32)
        d = cls(a.shape, dtype, zero=False)
        host = np.ascontiguousarray(a) if dtype == F32 else bf16_bits(a)
        if d.nbytes:
            d._ok(d.L.qsf_dev_copy(d.ptr, host.ctypes.data, d.nbytes, 1), "qsf_dev_copy")
        return d

    @classmethod
    def from_raw(cls, a, dtype):
        """An array uploaded byte for byte: uint8 (U8), int32 (I32), float32 (F32) or
        bfloat16 patterns as uint16 (BF16)."""
        np_type = {U8: np.uint8, I32: np.int32, F32: np.float32, BF16: np.uint16}[dtype]
        a = np.ascontiguousarray(a, np_type)
        d = cls(a.shape, dtype, zero=False)
        if d.nbytes:
            d._ok(d.L.qsf_dev_copy(d.ptr, a.ctypes.data, d.nbytes, 1), "qsf_dev_copy")
        return d

    @classmethod
    def from_bits(cls, bits):
        """A bfloat16 array from its bit patterns (uint16), uploaded as they are."""
        bits = np.ascontiguousarray(bits, np.uint16)
        d = cls(bits.shape, BF16, zero=False)
        if d.nbytes:
            d._ok(d.L.qsf_dev_copy(d.ptr, bits.ctypes.data, d.nbytes, 1), "qsf_dev_copy")
        return d

    def numpy(self) -> np.ndarray:
        host = np.empty(self.shape, np.float32 if self.dtype == F32 else np.uint16)
        if self.nbytes:
            self._ok(self.L.qsf_dev_copy(host.ctypes.data, self.ptr, self.nbytes, 2), "qsf_dev_copy")
        return host if self.dtype == F32 else bf16_to_f32(host)

    def bits(self) -> np.ndarray:
        """A bfloat16 array's bit patterns (uint16), as stored."""
        assert self.dtype == BF16
        host = np.empty(self.shape, np.uint16)
        if self.nbytes:
            self._ok(self.L.qsf_dev_copy(host.ctypes.data, self.ptr, self.nbytes, 2), "qsf_dev_copy")
        return host

    def zero_(self):
        if self.nbytes:
            self._ok(self.L.qsf_dev_zero(self.ptr, self.nbytes), "qsf_dev_zero")
        return self

    def item(self) -> float:
        return float(self.numpy().reshape(-1)[0])

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p:   # back to the pool: kernels queued on the null stream finish before any reuse
            _POOL.setdefault(self._block, []).append(p)
            self.ptr = None
