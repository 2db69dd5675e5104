"""Autoshape: the oracle pinned on the CPU; the HIP kernel against it on the GPU.

Golden cutoffs come from the reference's own tests: Coherent(0, alpha=1) ->
8 at the default probability 0.99999 (test_ket.py:98-100, test_dm.py:82-84),
Coherent(0, 1.0) -> 5 and SqueezedVacuum(r=+-0.5) -> 6 at 0.999
(test_branch_and_bound.py:40-54), Coherent(0, 0.0) -> 1 (test_coherent.py:75).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from mrmustard_amd import workloads as W
from tests.conftest import gpu_available

PROB, MAXS, MINS = 0.99999, 50, 1  # settings.py:67-74 defaults


def normalized_dm(A, b, c):
    """Scale c so that the trace is 1: Tr = c exp(-b^T (A - X)^-1 b / 2) / sqrt(det(A - X)),
    X = [[0, I], [I, 0]] (the reference's marginal formula applied to every mode)."""
    A = np.asarray(A, complex)
    b = np.asarray(b, complex)
    M = len(b) // 2
    X = np.block([[np.zeros((M, M)), np.eye(M)], [np.eye(M), np.zeros((M, M))]])
    R = A - X
    tr = c * np.exp(-0.5 * b @ np.linalg.solve(R, b)) / np.sqrt(np.linalg.det(R))
    return A, b, complex(c) / abs(tr)


def coherent_dm(alpha):
    return W.ket_to_dm(np.zeros((1, 1), complex), np.array([alpha], complex), np.exp(-abs(alpha) ** 2 / 2))


def squeezed_dm(r, phi=0.0):
    return W.ket_to_dm(np.array([[-np.exp(1j * phi) * np.tanh(r)]]), np.zeros(1, complex), 1 / np.sqrt(np.cosh(r)))


def first_cutoff(probs, max_prob, max_shape, min_shape):
    """Steps of the reference loop: the least k with sum_{n<=k} P(n) >= max_prob."""
    acc = 0.0
    for k, p in enumerate(probs):
        acc += p
        if acc >= max_prob:
            return int(np.clip(k, min_shape, max_shape))
    return max_shape


def random_dm(M, seed, loss=True):
    rng = np.random.default_rng(seed)
    A, b, c = W.gaussian_ket(M, rng, r_max=0.8, b_sigma=0.6)
    dm = W.ket_to_dm(A, b, c)
    if loss:
        dm = W.apply_loss(*dm, rng.uniform(0.6, 1.0, M))
    return normalized_dm(*dm)


# ----------------------------------------------------------------- oracle


def test_oracle_reference_cutoffs():
    from oracle import autoshape as O

    assert list(O.autoshape(*coherent_dm(1.0), PROB, MAXS, MINS)) == [8]
    assert list(O.autoshape(*coherent_dm(1.0), 0.999, MAXS, MINS)) == [5]
    for r in (0.5, -0.5):
        assert list(O.autoshape(*squeezed_dm(abs(r), 0.0 if r > 0 else np.pi), 0.999, MAXS, MINS)) == [6]
    assert list(O.autoshape(*coherent_dm(0.0), PROB, MAXS, MINS)) == [1]


@pytest.mark.parametrize("alpha", [0.3, 1.0, 2.0 + 1.0j, 3.5])
@pytest.mark.parametrize("p", [0.9, 0.999, 0.99999])
def test_oracle_poisson(alpha, p):
    """Coherent state: photon numbers are Poisson(|alpha|^2)."""
    from oracle import autoshape as O

    lam = abs(alpha) ** 2
    probs = [math.exp(-lam) * lam**n / math.factorial(n) for n in range(80)]
    assert O.autoshape(*coherent_dm(alpha), p, 60, 1)[0] == first_cutoff(probs, p, 60, 1)


@pytest.mark.parametrize("r", [0.2, 0.6, 1.0])
def test_oracle_squeezed(r):
    """Squeezed vacuum: P(2n) = tanh(r)^(2n) (2n)! / (2^n n!)^2 / cosh r, P(odd) = 0."""
    from oracle import autoshape as O

    t = math.tanh(r)
    probs = []
    for n in range(120):
        if n % 2:
            probs.append(0.0)
        else:
            h = n // 2
            probs.append(t ** n * math.factorial(n) / (2**h * math.factorial(h)) ** 2 / math.cosh(r))
    assert O.autoshape(*squeezed_dm(r), 0.999, 100, 1)[0] == first_cutoff(probs, 0.999, 100, 1)


def test_oracle_product_state_modes_independent():
    """A product of a coherent and a squeezed mode: each cutoff is its own."""
    from oracle import autoshape as O

    A = np.zeros((4, 4), complex)
    t = -np.tanh(0.6)
    A[1, 1] = np.conj(t)  # bra of mode 1
    A[3, 3] = t  # ket of mode 1
    b = np.array([np.conj(1.2), 0, 1.2, 0], complex)
    c = math.exp(-1.44) / math.cosh(0.6)
    got = O.autoshape(A, b, c, 0.999, 80, 1)
    assert got[0] == O.autoshape(*coherent_dm(1.2), 0.999, 80, 1)[0]
    assert got[1] == O.autoshape(*squeezed_dm(0.6), 0.999, 80, 1)[0]


@pytest.mark.parametrize("r,beta", [(0.4, (0.0, 0.0)), (0.7, (0.5, -0.3)), (0.9, (1.0, 0.4))])
def test_oracle_vs_fock_marginal(oracle_mod, r, beta):
    """Two-mode squeezed vacuum with a real linear term: the cutoffs equal
    those of the single-mode marginals (displaced thermal states) of its Fock
    density matrix, from the pinned forward oracle traced down to one mode.

    The reference's diagonal recursion (autoshape.py:130-152) keeps G[k+1, k]
    in the buffer slot that its next step reads as G[k, k+1] (and the like
    around the diagonal).  The two agree when the marginal's Fock matrix is
    symmetric (real coefficients, as here), when its A vanishes (coherent
    marginals) or when its off-diagonal A_01 does (pure squeezed modes): the
    closed-form tests above.  Other marginals follow the reference's formula
    as written, which the oracle restates and the kernel reproduces."""
    from oracle import autoshape as O

    t = np.tanh(r)
    A = np.array([[0, -t], [-t, 0]], complex)
    b = np.array(beta, complex)
    dm = normalized_dm(*W.ket_to_dm(A, b, 1.0))
    K = 60
    rho = oracle_mod.vanilla((K, K, K, K), *dm)  # [bra0, bra1, ket0, ket1]
    assert abs(np.einsum("ijij->", rho) - 1) < 1e-9
    p0 = np.real(np.einsum("ijkj->ik", rho).diagonal())
    p1 = np.real(np.einsum("ijil->jl", rho).diagonal())
    want = [first_cutoff(p0, 0.999, 50, 1), first_cutoff(p1, 0.999, 50, 1)]
    assert list(O.autoshape(*dm, 0.999, 50, 1)) == want


# ----------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
def test_autoshape_hip_reference_cutoffs():
    from mrmustard_amd.lattice.autoshape import autoshape_hip

    assert list(autoshape_hip(*coherent_dm(1.0), PROB, MAXS, MINS)) == [8]
    assert list(autoshape_hip(*coherent_dm(1.0), 0.999, MAXS, MINS)) == [5]
    assert list(autoshape_hip(*squeezed_dm(0.5), 0.999, MAXS, MINS)) == [6]
    assert list(autoshape_hip(*squeezed_dm(0.5, np.pi), 0.999, MAXS, MINS)) == [6]
    assert list(autoshape_hip(*coherent_dm(0.0), PROB, MAXS, MINS)) == [1]


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
@pytest.mark.parametrize("M", [1, 2, 3, 4, 8])
def test_autoshape_hip_vs_oracle(M):
    """Random lossy Gaussian states: the kernel's cutoffs equal the oracle's
    wherever the captured probability is not within 1e-9 of the threshold."""
    from oracle import autoshape as O

    from mrmustard_amd.lattice.autoshape import autoshape_batch_hip, autoshape_hip

    dms = [random_dm(M, 100 * M + s) for s in range(6)]
    for p, mx in ((0.99, 30), (0.99999, 50)):
        for A, b, c in dms:
            want = O.autoshape(A, b, c, p, mx, 1)
            got = autoshape_hip(A, b, c, p, mx, 1)
            sure = O.margins(A, b, c, p, mx) > 1e-9
            assert np.array_equal(got[sure], want[sure]), (M, p, got, want)
        got_b = autoshape_batch_hip(np.stack([d[0] for d in dms]), np.stack([d[1] for d in dms]),
                                    np.array([d[2] for d in dms]), p, mx, 1)
        for k, (A, b, c) in enumerate(dms):
            assert np.array_equal(got_b[k], autoshape_hip(A, b, c, p, mx, 1))


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
def test_autoshape_hip_errors():
    from mrmustard_amd.lattice.autoshape import autoshape_hip

    with pytest.raises(ValueError):
        autoshape_hip(np.zeros((3, 3), complex), np.zeros(3, complex), 1.0, 0.9, 10, 1)  # odd size
    with pytest.raises(ValueError):
        autoshape_hip(*coherent_dm(1.0), 0.9, 100000, 1)  # beyond the sqrt table
