"""Work-queue batches of triples with uneven mode counts and shapes.

The reference batches only triples that share one shape
(``vanilla_batch_numba``, vanilla/batch.py:27-61, a ``prange`` loop); a set of
triples with different D or cutoffs is a Python loop of ``vanilla_numba``
calls there.  ``HermitePlan`` covers both with one persistent-kernel launch:
the C ABI sorts the triples largest-first into a device work queue
(``mm_plan_create``), workgroups take the large ones, single waves the small
ones, single lanes the one-index chains; ``run`` launches it, optionally as a
captured HIP graph.

Packed layouts (complex128, triples in the given order): A and dA
concatenate D_k*D_k blocks, b and db D_k blocks, c and dc one entry each,
G and dLdG the prod(shape_k) amplitudes of each triple.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .lattice.strategies import C128, _ptr, _stream, as_device_c128


class HermitePlan:
    """A compiled work queue for a fixed list of shapes."""

    def __init__(self, shapes, fused: bool = False, tune: dict | None = None):
        L = _lib.load()
        self.shapes = [tuple(int(s) for s in sh) for sh in shapes]
        self.fused = bool(fused)
        n = len(self.shapes)
        self.dims = np.array([len(s) for s in self.shapes], dtype=np.int32)
        shp = np.ones((max(n, 1), 8), dtype=np.int64)
        for k, s in enumerate(self.shapes):
            if not 1 <= len(s) <= 8:
                raise ValueError(f"triple {k}: {len(s)} indices, expected 1..8")
            shp[k, : len(s)] = s
        t = _lib.MMTune()
        tune = dict(tune or {})
        t.block_threads = int(tune.get("block_threads", 0))
        t.blocks_per_cu = int(tune.get("blocks_per_cu", 0))
        t.wave_max_slab = int(tune.get("wave_max_slab", 0))
        self.tune = tune
        handle = ctypes.c_void_p()
        rc = L.mm_plan_create(n, self.dims.ctypes.data, shp.ctypes.data, int(self.fused), ctypes.byref(t), ctypes.byref(handle))
        _lib.check(rc, "mm_plan_create")
        self._h = handle
        na, ea, eb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(L.mm_plan_info(self._h, ctypes.byref(na), ctypes.byref(ea), ctypes.byref(eb)), "mm_plan_info")
        self.n_amplitudes, self.a_elems, self.b_elems = na.value, ea.value, eb.value
        sizes = np.array([int(np.prod(s)) for s in self.shapes], dtype=np.int64)
        self.g_off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64) if n else np.zeros(0, np.int64)
        d2 = self.dims.astype(np.int64)
        self.a_off = np.concatenate([[0], np.cumsum(d2 * d2)[:-1]]).astype(np.int64) if n else np.zeros(0, np.int64)
        self.b_off = np.concatenate([[0], np.cumsum(d2)[:-1]]).astype(np.int64) if n else np.zeros(0, np.int64)
        self.sizes = sizes

    def describe(self) -> dict:
        out = np.zeros(8, dtype=np.int64)
        _lib.check(_lib.load().mm_plan_describe(self._h, out.ctypes.data), "mm_plan_describe")
        keys = ["block_triples", "wave_triples", "lane_triples", "outside_triples", "grid", "threads", "lds_bytes", "wave_lds_bytes"]
        return dict(zip(keys, (int(x) for x in out)))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.load().mm_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------- packing
    def pack(self, triples):
        """Concatenate per-triple (A, b, c) into the plan's device layout."""
        if len(triples) != len(self.shapes):
            raise ValueError("one (A, b, c) per shape expected")
        A = np.concatenate([np.asarray(t[0], np.complex128).reshape(-1) for t in triples]) if triples else np.zeros(0, complex)
        b = np.concatenate([np.asarray(t[1], np.complex128).reshape(-1) for t in triples]) if triples else np.zeros(0, complex)
        c = np.array([complex(np.asarray(t[2]).reshape(())) for t in triples], dtype=np.complex128)
        if A.size != self.a_elems or b.size != self.b_elems:
            raise ValueError("triple sizes do not match the plan's shapes")
        return as_device_c128(A), as_device_c128(b), as_device_c128(c)

    def pack_amplitudes(self, arrays) -> torch.Tensor:
        """Concatenate per-triple amplitude arrays (e.g. cotangents) in plan order."""
        flat = [np.asarray(x, np.complex128).reshape(-1) for x in arrays]
        return as_device_c128(np.concatenate(flat) if flat else np.zeros(0, complex))

    def split(self, G: torch.Tensor):
        """Per-triple views of a packed amplitude buffer."""
        return [G[o : o + s].view(*sh) for o, s, sh in zip(self.g_off.tolist(), self.sizes.tolist(), self.shapes)]

    def split_grads(self, dA, db, dc):
        out = []
        for k, D in enumerate(self.dims.tolist()):
            a0, b0 = int(self.a_off[k]), int(self.b_off[k])
            out.append((dA[a0 : a0 + D * D].view(D, D), db[b0 : b0 + D], dc[k]))
        return out

    # ----------------------------------------------------------------- run
    def run(self, A, b, c, dLdG=None, G=None, grads=None, graph: bool = False):
        """Launch the plan on the current stream.

        Returns G (packed), or (G, dA, db, dc) for a fused plan.  ``G`` and
        ``grads`` may be given to reuse output buffers (needed for graph replay
        with stable pointers)."""
        L = _lib.load()
        dev = A.device
        if G is None:
            G = torch.empty(self.n_amplitudes, dtype=C128, device=dev)
        if self.fused:
            if dLdG is None:
                raise ValueError("fused plan needs dLdG")
            if grads is None:
                grads = (
                    torch.empty(self.a_elems, dtype=C128, device=dev),
                    torch.empty(self.b_elems, dtype=C128, device=dev),
                    torch.empty(len(self.shapes), dtype=C128, device=dev),
                )
            dA, db, dc = grads
        else:
            if dLdG is not None:
                raise ValueError("forward plan takes no dLdG")
            dA = db = dc = None
        fn = L.mm_plan_run_graph if graph else L.mm_plan_run
        rc = fn(self._h, _ptr(A), _ptr(b), _ptr(c), _ptr(dLdG), _ptr(G), _ptr(dA), _ptr(db), _ptr(dc), _stream())
        _lib.check(rc, "mm_plan_run_graph" if graph else "mm_plan_run")
        return (G, dA, db, dc) if self.fused else G
