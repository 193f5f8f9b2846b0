"""Minimal device-memory plumbing over the C ABI (no PyTorch needed)."""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import _lib


class _BlockCache:
    """Freed device blocks of pooled arrays kept for reuse, by size class (powers of two from 256
    bytes), up to a byte limit: hipMalloc / hipFree of per-request arrays would otherwise sit on
    the serving loop's critical path (hipFree waits for the device).  A pooled array must only be
    dropped when no queued GPU work still uses it (hipFree's implicit wait is skipped)."""

    LIMIT = 8 << 30

    def __init__(self):
        self.free_blocks = {}
        self.held = 0
        self.lock = threading.Lock()  # the serving loop frees while its admission thread allocates

    @staticmethod
    def size_class(nbytes: int) -> int:
        return max(256, 1 << (max(int(nbytes), 1) - 1).bit_length())

    def take(self, nbytes: int):
        with self.lock:
            blocks = self.free_blocks.get(self.size_class(nbytes))
            if blocks:
                p = blocks.pop()
                self.held -= self.size_class(nbytes)
                return p
        return None

    def give(self, ptr: C.c_void_p, cls: int) -> bool:
        with self.lock:
            if self.held + cls > self.LIMIT:
                return False
            self.free_blocks.setdefault(cls, []).append(ptr)
            self.held += cls
            return True


_cache = _BlockCache()


class PinnedBlock:
    """A pinned (page-locked) host buffer of a size class, from a PinnedPool."""

    def __init__(self, ptr: C.c_void_p, cls: int):
        self.ptr = ptr
        self.cls = cls


class PinnedPool:
    """Pinned host staging buffers for asynchronous uploads, reused by size class.  A block is
    taken for one upload and released once the stream has passed that copy (hipHostMalloc, so
    the copy engine reads it while the host moves on)."""

    def __init__(self):
        self.free_blocks = {}

    def take(self, nbytes: int) -> PinnedBlock:
        cls = _BlockCache.size_class(nbytes)
        blocks = self.free_blocks.get(cls)
        if blocks:
            return blocks.pop()
        p = C.c_void_p()
        _lib.check(_lib.load().esp_host_malloc_pinned(C.byref(p), cls))
        return PinnedBlock(p, cls)

    def release(self, block: PinnedBlock) -> None:
        self.free_blocks.setdefault(block.cls, []).append(block)

    def close(self) -> None:
        for blocks in self.free_blocks.values():
            for b in blocks:
                _lib.load().esp_host_free_pinned(b.ptr)
        self.free_blocks = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A device allocation with a numpy dtype and shape.  `pooled` arrays take and return their
    blocks through a bounded size-class cache (see _BlockCache)."""

    def __init__(self, shape, dtype=np.float64, zero: bool = False, pooled: bool = False):
        lib = _lib.load()
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._cls = _BlockCache.size_class(max(self.nbytes, 8)) if pooled else 0
        p = _cache.take(self._cls) if pooled else None
        if p is None:
            p = C.c_void_p()
            _lib.check(lib.esp_device_malloc(C.byref(p), self._cls if pooled else max(self.nbytes, 8)))
        self._ptr = p
        if zero:
            _lib.check(lib.esp_memset(self._ptr, 0, max(self.nbytes, 8)))

    @classmethod
    def from_numpy(cls, a: np.ndarray, pooled: bool = False) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, pooled=pooled)
        d.upload(a)
        return d

    @property
    def ptr(self) -> int:
        return self._ptr.value

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"size mismatch: {a.nbytes} != {self.nbytes}")
        if self.nbytes:
            _lib.check(_lib.load().esp_memcpy_htod(self._ptr, a.ctypes.data_as(C.c_void_p), self.nbytes))

    def upload_async(self, a: np.ndarray, stream: "Stream", pool: PinnedPool) -> PinnedBlock:
        """Copy `a` into a pinned staging block and enqueue its upload on `stream`; returns the
        block, which the caller releases to the pool once the stream has passed the copy."""
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"size mismatch: {a.nbytes} != {self.nbytes}")
        blk = pool.take(max(self.nbytes, 8))
        if self.nbytes:
            C.memmove(blk.ptr, a.ctypes.data, self.nbytes)
            _lib.check(_lib.load().esp_memcpy_htod_async(self._ptr, blk.ptr, self.nbytes, stream.handle))
        return blk

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _lib.check(_lib.load().esp_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), self._ptr, self.nbytes))
        return out

    def zero(self) -> None:
        if self.nbytes:
            _lib.check(_lib.load().esp_memset(self._ptr, 0, self.nbytes))

    def free(self) -> None:
        if getattr(self, "_ptr", None) is not None and self._ptr.value:
            if not (self._cls and _cache.give(self._ptr, self._cls)):
                _lib.load().esp_device_free(self._ptr)
        self._ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p()
        _lib.check(_lib.load().esp_stream_create(C.byref(s)))
        self.handle = s

    def synchronize(self) -> None:
        _lib.check(_lib.load().esp_stream_synchronize(self.handle))

    def __del__(self):
        try:
            if self.handle.value:
                _lib.load().esp_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        _lib.check(_lib.load().esp_event_create(C.byref(e)))
        self.handle = e

    def record(self, stream: Stream | None = None) -> None:
        _lib.check(_lib.load().esp_event_record(self.handle, stream.handle if stream else None))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        _lib.check(_lib.load().esp_event_elapsed_ms(self.handle, end.handle, C.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle.value:
                _lib.load().esp_event_destroy(self.handle)
        except Exception:
            pass


def synchronize() -> None:
    _lib.check(_lib.load().esp_device_synchronize())
