"""CPU oracle for the Bargmann->Fock hot path.  TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker.  The product
(``mrmustard_amd``) never imports it and fails loudly without its HIP library.

``hermite_oracle.c`` restates the reference kernels (file:line in its header):
  - ``vanilla``            <- mrmustard/math/lattice/strategies/vanilla/core.py:64-124
  - ``stable``             <- core.py:164-213
  - ``vanilla_vjp``        <- vanilla/gradients.py:41-82
  - ``vanilla_batch``      <- vanilla/batch.py:54-61
  - ``vanilla_batch_vjp``  <- vanilla/gradients.py:104-116
How each is pinned (closed forms, generating-function brute force, finite
differences, box independence) is listed in ``tests/test_oracle.py``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_i64p = ctypes.POINTER(ctypes.c_int64)
_dp = ctypes.c_void_p


def build() -> str:
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _load(_LIB_PATH)
    return _lib


_native = None


def host_tuned_lib():
    """The oracle compiled with -march=native for this host, built afresh into
    a temporary directory (a copy built elsewhere could use instructions this
    host lacks).  Same results as ``lib()``: no contraction, no fast math."""
    global _native
    if _native is None:
        import tempfile

        out = os.path.join(tempfile.mkdtemp(prefix="mm_oracle_"), "liboracle_native.so")
        subprocess.run(["make", "-s", "-C", _HERE, "native", f"NATIVE_OUT={out}"], check=True)
        _native = _load(out)
    return _native


def _load(path):
    L = ctypes.CDLL(path)
    for name in ("oracle_vanilla", "oracle_stable"):
        f = getattr(L, name)
        f.argtypes = [_i64p, ctypes.c_int, _dp, _dp, _dp, _dp]
        f.restype = ctypes.c_int
    for name in ("oracle_vjp", "oracle_vjp_kahan", "oracle_vjp_kahan_par"):
        f = getattr(L, name)
        f.argtypes = [_i64p, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        f.restype = ctypes.c_int
    L.oracle_vanilla_batch.argtypes = [_i64p, ctypes.c_int, ctypes.c_int64, _dp, _dp, _dp, ctypes.c_int, _dp]
    L.oracle_vanilla_batch.restype = ctypes.c_int
    L.oracle_vjp_batch.argtypes = [_i64p, ctypes.c_int, ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, _dp]
    L.oracle_vjp_batch.restype = ctypes.c_int
    L.oracle_uneven.argtypes = [ctypes.c_int64] + [_dp] * 13
    L.oracle_uneven.restype = ctypes.c_int
    L.oracle_max_threads.argtypes = []
    L.oracle_max_threads.restype = ctypes.c_int
    return L


def _c128(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.complex128))


def _shape_arr(shape) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(tuple(int(s) for s in shape), dtype=np.int64))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc: int, what: str):
    if rc != 0:
        raise ValueError(f"oracle {what} rejected its input (rc={rc})")


def vanilla(shape, A, b, c, out: np.ndarray | None = None) -> np.ndarray:
    """Oracle for ``vanilla_numba(shape, A, b, c, out)`` (core.py:25-124)."""
    shape = tuple(int(s) for s in shape)
    A, b, c = _c128(A), _c128(b), _c128(c).reshape(())
    D = b.shape[-1]
    G = out if out is not None else np.zeros(shape, dtype=np.complex128)
    assert G.flags.c_contiguous and G.dtype == np.complex128
    s = _shape_arr(shape)
    _check(lib().oracle_vanilla(s.ctypes.data_as(_i64p), D, _ptr(A), _ptr(b), _ptr(c), _ptr(G)), "vanilla")
    return G


def stable(shape, A, b, c, out: np.ndarray | None = None) -> np.ndarray:
    """Oracle for ``stable_numba(shape, A, b, c, out)`` (core.py:127-213)."""
    shape = tuple(int(s) for s in shape)
    A, b, c = _c128(A), _c128(b), _c128(c).reshape(())
    D = b.shape[-1]
    G = out if out is not None else np.zeros(shape, dtype=np.complex128)
    s = _shape_arr(shape)
    _check(lib().oracle_stable(s.ctypes.data_as(_i64p), D, _ptr(A), _ptr(b), _ptr(c), _ptr(G)), "stable")
    return G


def vanilla_vjp(G, c, dLdG, kahan: bool = False, threads: bool = False):
    """Oracle for ``vanilla_vjp_numba(G, c, dLdG)`` (gradients.py:25-82).

    ``kahan=True`` sums the same terms with compensation instead of the
    reference's left-to-right order; with ``threads=True`` as well, the
    compensated sums run over OpenMP threads (chunk sums added with
    compensation, for large lattices).
    """
    G, g, c = _c128(G), _c128(dLdG), _c128(c).reshape(())
    D = G.ndim
    dA = np.zeros((D, D), np.complex128)
    db = np.zeros(D, np.complex128)
    dc = np.zeros((), np.complex128)
    s = _shape_arr(G.shape)
    f = (lib().oracle_vjp_kahan_par if threads else lib().oracle_vjp_kahan) if kahan else lib().oracle_vjp
    _check(f(s.ctypes.data_as(_i64p), D, _ptr(G), _ptr(c), _ptr(g), _ptr(dA), _ptr(db), _ptr(dc)), "vjp")
    return dA, db, dc[()]


def vanilla_batch(shape, A, b, c, stable: bool = False) -> np.ndarray:
    """Oracle for ``vanilla_batch_numba`` (batch.py:27-61), OpenMP over the batch."""
    shape = tuple(int(x) for x in shape)
    A, b, c = _c128(A), _c128(b), _c128(c).reshape(-1)
    B, D = b.shape[0], b.shape[-1]
    G = np.zeros((B,) + shape, np.complex128)
    s = _shape_arr(shape)
    _check(lib().oracle_vanilla_batch(s.ctypes.data_as(_i64p), D, B, _ptr(A), _ptr(b), _ptr(c), int(stable), _ptr(G)), "batch")
    return G


def vanilla_batch_vjp(G, c, dLdG):
    """Oracle for ``vanilla_batch_vjp_numba`` (gradients.py:85-116)."""
    G, g, c = _c128(G), _c128(dLdG), _c128(c).reshape(-1)
    B, D = G.shape[0], G.ndim - 1
    dA = np.zeros((B, D, D), np.complex128)
    db = np.zeros((B, D), np.complex128)
    dc = np.zeros(B, np.complex128)
    s = _shape_arr(G.shape[1:])
    _check(lib().oracle_vjp_batch(s.ctypes.data_as(_i64p), D, B, _ptr(G), _ptr(c), _ptr(g), _ptr(dA), _ptr(db), _ptr(dc)), "batch vjp")
    return dA, db, dc


def uneven(triples, shapes, dLdG=None, host_tuned: bool = False):
    """Per-triple forward (and VJP when ``dLdG`` is given) over triples of
    uneven mode count and shape, OpenMP across triples (``host_tuned``: the
    -march=native build).

    Returns ``(Gs, grads)`` with ``Gs`` a list of arrays and ``grads`` a list of
    ``(dA, db, dc)`` or ``None``.
    """
    n = len(triples)
    dims = np.array([len(s) for s in shapes], np.int32)
    shp = np.zeros((n, 8), np.int64)
    for k, s in enumerate(shapes):
        shp[k, : len(s)] = s
    sizes = np.array([int(np.prod(s)) for s in shapes], np.int64)
    g_off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    a_off = np.concatenate([[0], np.cumsum(dims.astype(np.int64) ** 2)[:-1]]).astype(np.int64)
    b_off = np.concatenate([[0], np.cumsum(dims.astype(np.int64))[:-1]]).astype(np.int64)
    A = np.concatenate([_c128(t[0]).ravel() for t in triples])
    b = np.concatenate([_c128(t[1]).ravel() for t in triples])
    c = np.array([complex(np.asarray(t[2]).reshape(())) for t in triples], np.complex128)
    G = np.zeros(int(sizes.sum()), np.complex128)
    g = None if dLdG is None else np.concatenate([_c128(x).ravel() for x in dLdG])
    dA = np.zeros_like(A)
    db = np.zeros_like(b)
    dc = np.zeros_like(c)
    rc = (host_tuned_lib() if host_tuned else lib()).oracle_uneven(
        n, _ptr(dims), _ptr(shp), _ptr(a_off), _ptr(b_off), _ptr(g_off), _ptr(A), _ptr(b), _ptr(c), _ptr(G),
        None if g is None else _ptr(g), _ptr(dA), _ptr(db), _ptr(dc),
    )
    _check(rc, "uneven")
    Gs = [G[o : o + s].reshape(shapes[k]) for k, (o, s) in enumerate(zip(g_off, sizes))]
    grads = None
    if g is not None:
        grads = [
            (dA[a_off[k] : a_off[k] + dims[k] ** 2].reshape(dims[k], dims[k]), db[b_off[k] : b_off[k] + dims[k]], dc[k])
            for k in range(n)
        ]
    return Gs, grads


def max_threads() -> int:
    return int(lib().oracle_max_threads())
