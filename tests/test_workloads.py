"""The synthetic workload generators (mrmustard_amd/workloads.py) and the
MrMustard shim's wiring (no GPU needed)."""

from __future__ import annotations

from math import factorial

import numpy as np

from mrmustard_amd import workloads as W


def test_haar_unitary_is_unitary():
    rng = np.random.default_rng(0)
    for m in (1, 2, 3, 4):
        U = W.haar_unitary(m, rng)
        assert np.allclose(U @ U.conj().T, np.eye(m), atol=1e-13)


def test_gaussian_ket_is_symmetric_and_contracting():
    rng = np.random.default_rng(1)
    A, b, c = W.gaussian_ket(3, rng)
    assert np.allclose(A, A.T)
    assert np.max(np.abs(np.linalg.eigvals(A))) < 1
    assert complex(c) == 1


def _kraus_loss(D, eta, cut, K):
    """Single-mode pure loss on a Fock density matrix D[bra, ket] via Kraus operators."""
    out = np.zeros((K, K), complex)
    for mp in range(K):
        for npp in range(K):
            for l in range(cut - max(mp, npp)):
                out[mp, npp] += (
                    (1 - eta) ** l / factorial(l) * eta ** ((mp + npp) / 2)
                    * np.sqrt(factorial(npp + l) / factorial(npp) * factorial(mp + l) / factorial(mp)) * D[mp + l, npp + l]
                )
    return out


def test_loss_map_matches_kraus_sum(oracle_mod):
    rng = np.random.default_rng(5)
    A, b, c = W.gaussian_ket(1, rng)
    cut, K = 30, 10
    psi = oracle_mod.vanilla((cut,), A, b, c)
    Dm = np.conj(psi)[:, None] * psi[None, :]
    eta = 0.73
    A2, b2, c2 = W.apply_loss(*W.ket_to_dm(A, b, c), [eta])
    G = oracle_mod.vanilla((K, K), A2, b2, c2)
    assert np.max(np.abs(G - _kraus_loss(Dm, eta, cut, K))) < 1e-9


def test_displacement_rule_matches_coherent_state(oracle_mod):
    """Displacing the vacuum gives the coherent-state triple (triples.py:103-123)."""
    alpha = np.array([0.4 - 0.3j])
    A, b, c = W.displace_ket(np.zeros((1, 1), complex), np.zeros(1, complex), 1.0, alpha)
    assert np.allclose(b, alpha) and np.isclose(c, np.exp(-0.5 * abs(alpha[0]) ** 2))


def test_uneven_capping_rule():
    triples, shapes = W.uneven_kets(n=500, seed=9)
    for (A, b, c), s in zip(triples, shapes):
        assert len(s) == A.shape[0] == b.shape[0]
        assert 1 <= len(s) <= 3
        assert int(np.prod(s)) <= 200_000
        assert all(1 <= x <= 60 for x in s)


def test_config_shapes():
    (A, b, c), shape = W.plain_ket_50x50()
    assert shape == (50, 50) and A.shape == (2, 2)
    (A, b, c), shape = W.lossy_dm_40x4()
    assert shape == (40,) * 4 and A.shape == (4, 4) and np.allclose(A, A.T)
    (A, b, c), shape, g = W.batched_kets_30x3(n=4)
    assert A.shape == (4, 3, 3) and g.shape == (4, 30, 30, 30)
    (A, b, c), shape = W.dm8_10x8()
    assert shape == (10,) * 8 and A.shape == (8, 8) and np.allclose(A, A.T)


def test_shim_wiring():
    from mrmustard_amd import shim

    class FakeNumpyBackend:
        name = "numpy"

        def other(self):
            return "forwarded"

    class FakeManager:
        _backend = FakeNumpyBackend()

    m = FakeManager()
    w = shim.install(m)
    assert m._backend is w and m._backend.other() == "forwarded" and m._backend.name == "numpy"
    assert shim.install(m) is w
    shim.uninstall(m)
    assert isinstance(m._backend, FakeNumpyBackend)
    # the numpy backend's Hermite methods (backend_numpy.py:381-446) are the wrapper's own
    for name in ("hermite_renormalized", "hermite_renormalized_batched", "hermite_renormalized_binomial",
                 "hermite_renormalized_diagonal", "hermite_renormalized_1leftoverMode"):
        assert name in vars(shim.HipBackendWrapper), name
