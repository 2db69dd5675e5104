"""Host-side checks that need no GPU: the C ABI library loads and exports
every symbol of include/mm_hermite.h, and the backend manager reproduces the
reference's dispatch (backend_manager.py:643-727, test_vanilla.py:165-238)."""

from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "mm_hermite.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    from mrmustard_amd import _lib

    path = _lib.LIB_PATH
    if not os.path.exists(path):
        pytest.skip("libmmhip.so not built")
    L = ctypes.CDLL(path)
    syms = header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the Python binding declares exactly the header's entry points
    assert sorted(_lib.SIGNATURES) == syms


def test_error_strings_without_gpu():
    from mrmustard_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmmhip.so not built")
    L = _lib.load()
    assert L.mm_version().decode().startswith("mmhip")
    assert isinstance(L.mm_last_error(), bytes)


def test_tune_struct_matches_header():
    from mrmustard_amd import _lib

    assert ctypes.sizeof(_lib.MMTune) == 16


# -------------------------------------------------------------- dispatch


class RecordingBackend:
    """Stands in for BackendHip: records which method the manager picks."""

    name = "rec"
    int32 = int64 = float32 = float64 = complex64 = complex128 = None

    def __init__(self):
        self.calls = []

    def astensor(self, x, dtype=None):
        return np.asarray(x)

    def reshape(self, x, shape):
        return np.reshape(x, shape)

    def broadcast_to(self, x, shape):
        return np.broadcast_to(x, shape)

    def hermite_renormalized(self, A, b, c, shape, stable=False, out=None):
        self.calls.append(("unbatched", A.shape, b.shape, np.shape(c), shape, stable, out))
        return np.zeros(shape, complex) if out is None else out

    def hermite_renormalized_batched(self, A, b, c, shape, stable=False, out=None):
        self.calls.append(("batched", A.shape, b.shape, np.shape(c), shape, stable, None if out is None else out.shape))
        return np.zeros((A.shape[0], *shape), complex)


@pytest.fixture
def manager():
    from mrmustard_amd.math.backend_manager import BackendManager

    m = BackendManager()
    saved = m._backend
    rec = RecordingBackend()
    m._backend = rec
    yield m, rec
    m._backend = saved


def test_singleton():
    from mrmustard_amd import math
    from mrmustard_amd.math.backend_manager import BackendManager

    assert BackendManager() is math


def test_unbatched_dispatch(manager):
    m, rec = manager
    G = m.hermite_renormalized(np.eye(2), np.ones(2), 1.0, (3, 3))
    assert G.shape == (3, 3)
    assert rec.calls[-1][0] == "unbatched" and rec.calls[-1][4] == (3, 3)


def test_b_batched_dispatch(manager):
    m, rec = manager
    G = m.hermite_renormalized(np.eye(2), np.ones((2, 1, 2)), 1.0, (4, 5))
    assert G.shape == (2, 1, 4, 5)
    kind, a_shape, b_shape, c_shape, shape, _, _ = rec.calls[-1]
    assert kind == "batched" and a_shape == (2, 2, 2) and b_shape == (2, 2) and c_shape == (2,)


def test_fully_batched_dispatch(manager):
    m, rec = manager
    G = m.hermite_renormalized(np.ones((2, 1, 2, 2)), np.ones((2, 1, 2)), np.ones((2, 1)), (4, 5))
    assert G.shape == (2, 1, 4, 5)
    kind, a_shape, b_shape, c_shape, *_ = rec.calls[-1]
    assert kind == "batched" and a_shape == (2, 2, 2) and b_shape == (2, 2) and c_shape == (2,)


def test_batch_shape_errors(manager):
    m, _ = manager
    with pytest.raises(ValueError, match="must match batch_shape"):
        m.hermite_renormalized(np.ones((2, 2, 2)), np.ones((3, 2)), np.ones(2), (3, 3))
    with pytest.raises(ValueError, match="must match batch_shape"):
        m.hermite_renormalized(np.ones((2, 2, 2)), np.ones((2, 2)), np.ones(3), (3, 3))


def test_out_too_small(manager):
    m, _ = manager
    with pytest.raises(ValueError, match="too large for out.shape"):
        m.hermite_renormalized(np.eye(2), np.ones(2), 1.0, (3, 3), out=np.zeros((2, 3), complex))
    with pytest.raises(ValueError, match="too large for out.shape"):
        m.hermite_renormalized(np.eye(2), np.ones((2, 2)), 1.0, (3, 3), out=np.zeros((2, 3, 2), complex))


def test_out_returned_unbatched(manager):
    m, _ = manager
    out = np.zeros((3, 3), complex)
    assert m.hermite_renormalized(np.eye(2), np.ones(2), 1.0, (3, 3), out=out) is out


def test_stable_flag_from_settings(manager):
    from mrmustard_amd import settings

    m, rec = manager
    m.hermite_renormalized(np.eye(2), np.ones(2), 1.0, (3, 3))
    assert rec.calls[-1][5] is False
    with settings(STABLE_FOCK_CONVERSION=True):
        m.hermite_renormalized(np.eye(2), np.ones(2), 1.0, (3, 3))
        assert rec.calls[-1][5] is True
    m.hermite_renormalized(np.eye(2), np.ones(2), 1.0, (3, 3))
    assert rec.calls[-1][5] is False


def test_missing_method_raises_not_implemented(manager):
    m, _ = manager
    with pytest.raises(NotImplementedError):
        m._apply("hermite_renormalized_binomial", ())


def test_unknown_backend():
    from mrmustard_amd import math

    with pytest.raises(ValueError):
        math.change_backend("numpy-that-does-not-exist")
    assert math.backend_name == "hip"


def test_cpu_tensor_inputs_fail_loudly_without_gpu():
    """Without a GPU the product path raises: there is no CPU fallback."""
    import torch

    from mrmustard_amd import math

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError):
        math.hermite_renormalized(np.eye(2) * 0.1, np.ones(2), 1.0, (3, 3))
