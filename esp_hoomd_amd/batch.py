"""Batched ESP solves: many independent systems, one launch per stage.

``Batch.enqueue(pairs)`` evaluates every ``(Solver, DeviceSystem)`` pair exactly as
``Solver.enqueue`` would, but with one kernel launch per stage for the whole set, captured
into a hipGraph and replayed while the set keeps its shape (same solvers, particle counts,
meshes and flags).
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence, Tuple

from . import _lib
from .solver import DeviceSystem, Solver


class Batch:
    def __init__(self, use_graph: bool = True, max_grid: int = 2048):
        b = C.c_void_p()
        _lib.check(_lib.load().esp_batch_create(C.byref(b)))
        self._b = b
        self.set_options(use_graph, max_grid)
        self._keep = None

    def set_options(self, use_graph: bool = True, max_grid: int = 2048) -> None:
        _lib.check_batch(_lib.load().esp_batch_set_options(self._b, 1 if use_graph else 0, int(max_grid)), self._b)

    def enqueue(self, pairs: Sequence[Tuple[Solver, DeviceSystem]], flags: int = 0, stream=None) -> None:
        n = len(pairs)
        ctxs = (C.c_void_p * max(n, 1))()
        systems = (_lib.EspSystem * max(n, 1))()
        outs = (_lib.EspOut * max(n, 1))()
        for i, (solver, system) in enumerate(pairs):
            ctxs[i] = solver.ctx.value
            systems[i] = system.c_system()
            outs[i] = system.c_out(flags)
        self._keep = (ctxs, systems, outs)
        _lib.check_batch(_lib.load().esp_batch_compute(self._b, n, ctxs, systems, outs,
                                                       stream.handle if stream is not None else None), self._b)

    def prepare(self, pairs: Sequence[Tuple[Solver, DeviceSystem]], flags: int = 0):
        """Build the argument arrays once for a fixed set of systems (enqueue_prepared reuses
        them; refresh(i) updates one system's entry, e.g. after its neighbour list moved)."""
        n = len(pairs)
        ctxs = (C.c_void_p * max(n, 1))()
        systems = (_lib.EspSystem * max(n, 1))()
        outs = (_lib.EspOut * max(n, 1))()
        for i, (solver, system) in enumerate(pairs):
            ctxs[i] = solver.ctx.value
            systems[i] = system.c_system()
            outs[i] = system.c_out(flags)
        return {"n": n, "ctxs": ctxs, "systems": systems, "outs": outs, "pairs": list(pairs)}

    @staticmethod
    def refresh(prepared, i: int) -> None:
        prepared["systems"][i] = prepared["pairs"][i][1].c_system()

    def enqueue_prepared(self, prepared, stream=None) -> None:
        _lib.check_batch(_lib.load().esp_batch_compute(self._b, prepared["n"], prepared["ctxs"], prepared["systems"],
                                                       prepared["outs"], stream.handle if stream is not None else None),
                         self._b)

    def compute(self, pairs: Sequence[Tuple[Solver, DeviceSystem]], flags: int = 0):
        """Evaluate and return host copies per system: force (N, 4), virial (6, N), ext (7,)."""
        self.enqueue(pairs, flags)
        _lib.check(_lib.load().esp_device_synchronize())
        return [{"force": s.force.numpy(), "virial": s.virial.numpy()[:, : s.N], "ext": s.ext.numpy()[:7]}
                for _, s in pairs]

    def stats(self) -> dict:
        c, r, l = C.c_ulonglong(), C.c_ulonglong(), C.c_ulonglong()
        _lib.check_batch(_lib.load().esp_batch_stats(self._b, C.byref(c), C.byref(r), C.byref(l)), self._b)
        return {"captures": c.value, "replays": r.value, "launches": l.value}

    def close(self) -> None:
        if getattr(self, "_b", None) is not None and self._b.value:
            _lib.load().esp_batch_destroy(self._b)
        self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Batch"]
