"""Pins the CPU oracle (oracle/hermite_oracle.c) before it is trusted.

Each check is independent of the recurrence it tests, and chosen so that a
plausible slip in the restatement (a dropped term, a transposed A, a wrong
sqrt index, a missing 1/2, a conjugated cotangent) fails at least one:

forward (core.py:64-124)
  - closed forms of the reference's own tests: Hermite polynomials
    (test_special.py:23-36), coherent (test_fock_utils.py:101-104), squeezed
    (:117-129), two-mode squeezed (:77), displacement matrix elements
    (the ``displacement`` strategy used in test_lattice_functions.py:141-147);
  - A = 0 and diagonal A (products of one-index sequences);
  - the generating function expanded as a power series in 50-digit mpmath
    (tests/brute.py), on random non-symmetric A (only the upper triangle
    counts, core.py:101-103) and shapes with size-1 axes;
  - box independence: a sub-box equals the block of a larger box, bit for bit.
stable (core.py:164-213)
  - the same brute force with symmetric A; the displacement closed form.
VJP (gradients.py:41-82)
  - central finite differences of the (pinned) forward, with the reference's
    conventions: no conjugation of dL/dG (test_vanilla.py:60) and dA compared
    with (fd + fd.T)/2 (test_vanilla.py:87);
  - linearity in dL/dG; the compensated variant agrees with the sequential one.
batch (batch.py:54-61, gradients.py:104-116) and uneven batches
  - equal, bit for bit, to per-triple calls.
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy.special import eval_genlaguerre, eval_hermite, factorial

from tests.brute import hermite_series, stable_tuple_recursion


def rel(x, ref):
    x, ref = np.asarray(x), np.asarray(ref)
    return float(np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300))


def random_triple(rng, D, sym=False, scale=0.45):
    A = rng.normal(size=(D, D)) + 1j * rng.normal(size=(D, D))
    if sym:
        A = A + A.T
    A *= scale / max(1.0, np.max(np.abs(np.linalg.eigvals(A))))
    b = rng.normal(size=D) + 1j * rng.normal(size=D)
    c = complex(rng.normal(), rng.normal())
    return A, b, c


# ------------------------------------------------------------- closed forms


def test_hermite_polynomials(oracle_mod):
    """test_special.py:23-36: A=[[-2]], b=[2x], c=1 gives H_n(x)/sqrt(n!)."""
    n = 12
    for x in np.linspace(-1.5, 1.5, 13):
        G = oracle_mod.vanilla((n,), np.array([[-2.0]]), np.array([2 * x]), 1.0)
        expected = np.array([eval_hermite(k, x) / np.sqrt(factorial(k)) for k in range(n)])
        assert np.allclose(G, expected, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("alpha", [0.0, 0.3 - 0.2j, 1.1 + 0.7j, -1.9j])
def test_coherent(oracle_mod, alpha):
    """test_fock_utils.py:101-104 (coherent_state_Abc: A=0, b=alpha)."""
    n = 30
    c = np.exp(-0.5 * abs(alpha) ** 2)
    G = oracle_mod.vanilla((n,), np.zeros((1, 1)), np.array([alpha]), c)
    expected = c * np.array([alpha**k / np.sqrt(factorial(k)) for k in range(n)])
    assert rel(G, expected) < 1e-13


@pytest.mark.parametrize("r,phi", [(0.4, 0.0), (1.3, 0.7), (2.0, -2.1)])
def test_squeezed(oracle_mod, r, phi):
    """test_fock_utils.py:117-129 (squeezed_vacuum_state_Abc, triples.py:126-150)."""
    n = 40
    A = np.array([[-np.tanh(r) * np.exp(1j * phi)]])
    G = oracle_mod.vanilla((n,), A, np.zeros(1), 1 / np.sqrt(np.cosh(r)))
    assert np.all(G[1::2] == 0)
    m = np.arange(n // 2)
    expected = (-np.exp(1j * phi) * np.tanh(r)) ** m * np.sqrt(factorial(2 * m)) / (2.0**m * factorial(m)) / np.sqrt(np.cosh(r))
    assert rel(G[0::2], expected) < 1e-12


@pytest.mark.parametrize("r,phi", [(0.3, 0.2), (1.0, -1.0)])
def test_two_mode_squeezed(oracle_mod, r, phi):
    """test_fock_utils.py:70-78 (two_mode_squeezed_vacuum_state_Abc, triples.py:193-221)."""
    n = 20
    t = np.exp(1j * phi) * np.tanh(r)
    A = np.array([[0, t], [t, 0]])
    G = oracle_mod.vanilla((n, n), A, np.zeros(2), 1 / np.cosh(r))
    expected = np.diag((1 / np.cosh(r)) * t ** np.arange(n))
    assert rel(G, expected) < 1e-13


def displacement_matrix(alpha, n):
    x = abs(alpha) ** 2
    E = np.zeros((n, n), complex)
    for m in range(n):
        for k in range(n):
            if m >= k:
                E[m, k] = np.sqrt(factorial(k) / factorial(m)) * alpha ** (m - k) * np.exp(-x / 2) * eval_genlaguerre(k, m - k, x)
            else:
                E[m, k] = np.sqrt(factorial(m) / factorial(k)) * (-np.conj(alpha)) ** (k - m) * np.exp(-x / 2) * eval_genlaguerre(m, k - m, x)
    return E


@pytest.mark.parametrize("which", ["vanilla", "stable"])
def test_displacement_gate(oracle_mod, which):
    """<m|D(alpha)|n> from displacement_gate_Abc (triples.py:424-446)."""
    alpha = 0.7 + 0.4j
    A = np.array([[0, 1], [1, 0]], complex)
    b = np.array([alpha, -np.conj(alpha)])
    c = np.exp(-abs(alpha) ** 2 / 2)
    G = getattr(oracle_mod, which)((16, 16), A, b, c)
    assert rel(G, displacement_matrix(alpha, 16)) < 1e-13


def test_zero_A_is_product(oracle_mod):
    rng = np.random.default_rng(1)
    b = rng.normal(size=3) + 1j * rng.normal(size=3)
    c = 0.3 - 0.8j
    shape = (5, 4, 6)
    G = oracle_mod.vanilla(shape, np.zeros((3, 3)), b, c)
    one = [np.array([b[i] ** k / np.sqrt(factorial(k)) for k in range(shape[i])]) for i in range(3)]
    expected = c * np.einsum("i,j,k->ijk", *one)
    assert rel(G, expected) < 1e-14


def test_diagonal_A_is_product(oracle_mod):
    rng = np.random.default_rng(2)
    a = 0.6 * (rng.normal(size=3) + 1j * rng.normal(size=3)) / 2
    b = rng.normal(size=3) + 1j * rng.normal(size=3)
    shape = (7, 3, 5)
    G = oracle_mod.vanilla(shape, np.diag(a), b, 1.0)
    one = [oracle_mod.vanilla((shape[i],), np.array([[a[i]]]), np.array([b[i]]), 1.0) for i in range(3)]
    assert rel(G, np.einsum("i,j,k->ijk", *one)) < 1e-14


# -------------------------------------------------------------- brute force

SHAPES = [(9,), (5, 4), (1, 6), (6, 1), (3, 4, 2), (2, 1, 3), (2, 3, 2, 2)]


@pytest.mark.parametrize("shape", SHAPES)
def test_vanilla_vs_power_series(oracle_mod, shape):
    rng = np.random.default_rng(len(shape) * 100 + sum(shape))
    A, b, c = random_triple(rng, len(shape), sym=False)
    G = oracle_mod.vanilla(shape, A, b, c)
    ref = hermite_series(A, b, c, shape, upper_only=True)
    assert rel(G, ref) < 1e-13
    # a transposed A (lower triangle used) must be detected
    if len(shape) > 1 and min(shape) > 1:
        assert rel(hermite_series(A.T, b, c, shape, upper_only=True), ref) > 1e-6


@pytest.mark.parametrize("shape", SHAPES)
def test_stable_vs_power_series(oracle_mod, shape):
    rng = np.random.default_rng(len(shape) * 77 + sum(shape))
    A, b, c = random_triple(rng, len(shape), sym=True)
    G = oracle_mod.stable(shape, A, b, c)
    assert rel(G, hermite_series(A, b, c, shape, upper_only=True)) < 1e-13


@pytest.mark.parametrize("shape", [(2, 2), (3, 3), (4, 3), (2, 3, 2), (3, 2, 2, 2)])
def test_stable_nonsymmetric_vs_tuple_recursion(oracle_mod, shape):
    """On a non-symmetric A the all-pivot average reads both triangles, and the
    pivots' recurrences disagree, so no generating function applies: the
    oracle is checked against an independent tuple-indexed recursion in 50-digit
    arithmetic (tests/brute.py).  A transposed A (A[j, i] read for A[i, j])
    must be detected; on a (2, 2) box it cannot be, since the only point with
    two pivots averages A[0, 1] and A[1, 0]."""
    rng = np.random.default_rng(len(shape) * 31 + sum(shape))
    A, b, c = random_triple(rng, len(shape), sym=False)
    assert np.max(np.abs(A - A.T)) > 0.1
    G = oracle_mod.stable(shape, A, b, c)
    ref = stable_tuple_recursion(A, b, c, shape)
    assert rel(G, ref) < 1e-13
    if max(shape) > 2:
        assert rel(stable_tuple_recursion(A.T, b, c, shape), ref) > 1e-6
        assert rel(oracle_mod.stable(shape, A.T, b, c), ref) > 1e-6


def test_box_independence(oracle_mod):
    rng = np.random.default_rng(5)
    A, b, c = random_triple(rng, 3)
    big = oracle_mod.vanilla((9, 7, 8), A, b, c)
    small = oracle_mod.vanilla((4, 5, 3), A, b, c)
    assert np.array_equal(small, big[:4, :5, :3])


def test_out_buffer_is_written_in_place(oracle_mod):
    rng = np.random.default_rng(6)
    A, b, c = random_triple(rng, 2)
    out = np.zeros((5, 6), np.complex128)
    G = oracle_mod.vanilla((5, 6), A, b, c, out=out)
    assert G is out and np.array_equal(out, oracle_mod.vanilla((5, 6), A, b, c))


# ---------------------------------------------------------------------- VJP


def fd_vjp(oracle_mod, A, b, c, shape, g, h=1e-6):
    """Central differences of sum(g * G) along the real direction of each parameter."""

    def L(A_, b_, c_):
        return np.sum(g * oracle_mod.vanilla(shape, A_, b_, c_))

    D = len(shape)
    dA = np.zeros((D, D), complex)
    for i in range(D):
        for j in range(D):
            E = np.zeros((D, D))
            E[i, j] = h
            dA[i, j] = (L(A + E, b, c) - L(A - E, b, c)) / (2 * h)
    db = np.zeros(D, complex)
    for i in range(D):
        e = np.zeros(D)
        e[i] = h
        db[i] = (L(A, b + e, c) - L(A, b - e, c)) / (2 * h)
    dc = (L(A, b, c + h) - L(A, b, c - h)) / (2 * h)
    return dA, db, dc


@pytest.mark.parametrize("shape", [(7,), (4, 4), (4, 3, 5), (1, 2, 3), (3, 2, 2, 3)])
def test_vjp_vs_finite_differences(oracle_mod, shape):
    rng = np.random.default_rng(673 + len(shape))
    A, b, c = random_triple(rng, len(shape))
    G = oracle_mod.vanilla(shape, A, b, c)
    g = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    dA, db, dc = oracle_mod.vanilla_vjp(G, c, g)
    fA, fb, fc = fd_vjp(oracle_mod, A, b, c, shape, g)
    assert abs(dc - fc) <= 1e-7 * abs(fc)
    assert rel(db, fb) < 1e-7
    assert rel(dA, (fA + fA.T) / 2) < 1e-7
    # a conjugated cotangent would be caught
    cA, cb, cc = oracle_mod.vanilla_vjp(G, c, np.conj(g))
    assert rel(cb, fb) > 1e-3


def test_vjp_linear_in_cotangent(oracle_mod):
    rng = np.random.default_rng(11)
    A, b, c = random_triple(rng, 3)
    shape = (5, 6, 4)
    G = oracle_mod.vanilla(shape, A, b, c)
    g1 = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    g2 = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    r1 = oracle_mod.vanilla_vjp(G, c, g1)
    r2 = oracle_mod.vanilla_vjp(G, c, g2)
    r12 = oracle_mod.vanilla_vjp(G, c, g1 + 2j * g2)
    for x1, x2, x12 in zip(r1, r2, r12):
        assert rel(x12, np.asarray(x1) + 2j * np.asarray(x2)) < 1e-12


def test_vjp_kahan_matches_sequential(oracle_mod):
    rng = np.random.default_rng(12)
    A, b, c = random_triple(rng, 3)
    shape = (12, 11, 10)
    G = oracle_mod.vanilla(shape, A, b, c)
    g = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    for x, y in zip(oracle_mod.vanilla_vjp(G, c, g), oracle_mod.vanilla_vjp(G, c, g, kahan=True)):
        assert rel(x, y) < 1e-13


# ------------------------------------------------------------------ batches


def test_batch_equals_loop(oracle_mod):
    rng = np.random.default_rng(13)
    B, shape = 5, (4, 6, 3)
    trip = [random_triple(rng, 3) for _ in range(B)]
    A = np.stack([t[0] for t in trip])
    b = np.stack([t[1] for t in trip])
    c = np.array([t[2] for t in trip])
    for stable in (False, True):
        G = oracle_mod.vanilla_batch(shape, A, b, c, stable=stable)
        one = oracle_mod.stable if stable else oracle_mod.vanilla
        for k in range(B):
            assert np.array_equal(G[k], one(shape, A[k], b[k], c[k]))
    g = rng.normal(size=G.shape) + 1j * rng.normal(size=G.shape)
    G = oracle_mod.vanilla_batch(shape, A, b, c)
    dA, db, dc = oracle_mod.vanilla_batch_vjp(G, c, g)
    for k in range(B):
        rA, rb, rc = oracle_mod.vanilla_vjp(G[k], c[k], g[k])
        assert np.array_equal(dA[k], rA) and np.array_equal(db[k], rb) and dc[k] == rc


def test_uneven_equals_loop(oracle_mod):
    rng = np.random.default_rng(14)
    shapes = [(7,), (3, 5), (2, 3, 4), (6, 2), (1,), (4, 1, 2, 3)]
    trip = [random_triple(rng, len(s)) for s in shapes]
    gs = [rng.normal(size=s) + 1j * rng.normal(size=s) for s in shapes]
    Gs, grads = oracle_mod.uneven(trip, shapes, gs)
    for (A, b, c), s, G, g, (dA, db, dc) in zip(trip, shapes, Gs, gs, grads):
        ref = oracle_mod.vanilla(s, A, b, c)
        assert np.array_equal(G, ref)
        rA, rb, rc = oracle_mod.vanilla_vjp(ref, c, g)
        assert np.array_equal(dA, rA) and np.array_equal(db, rb) and dc == rc


# ------------------------------------------------------- displacement oracle


def _displacement_closed_form(alpha, n, m):
    """<n|D(alpha)|m> in 50-digit arithmetic from the generalized Laguerre
    closed form (Cahill and Glauber, Phys. Rev. 177, 1857 (1969))."""
    import mpmath

    mpmath.mp.dps = 50
    al = mpmath.mpc(alpha.real, alpha.imag)
    x = abs(al) ** 2
    if n >= m:
        v = mpmath.sqrt(mpmath.factorial(m) / mpmath.factorial(n)) * al ** (n - m) * mpmath.laguerre(m, n - m, x)
    else:
        v = mpmath.sqrt(mpmath.factorial(n) / mpmath.factorial(m)) * (-mpmath.conj(al)) ** (m - n) * mpmath.laguerre(n, m - n, x)
    return complex(v * mpmath.exp(-x / 2))


@pytest.mark.parametrize("alpha", [4.0 + 4.0j, 0.3 - 1.1j, -2.5 + 0.0j])
@pytest.mark.parametrize("cutoffs", [(150, 150), (150, 90), (60, 150)])
def test_displacement_vs_laguerre_closed_form(alpha, cutoffs):
    from oracle.displacement import displacement

    D = displacement(cutoffs, alpha)
    assert D.shape == cutoffs
    rows, cols = range(0, cutoffs[0], 3), range(1, cutoffs[1], 3)
    ref = np.array([[_displacement_closed_form(alpha, n, m) for m in cols] for n in rows])
    assert rel(D[::3, 1::3], ref) < 1e-12
    # a conjugated phase (the sign of i (n - m) arg(alpha)) is detected
    if alpha.imag != 0:
        assert rel(displacement(cutoffs, np.conj(alpha))[::3, 1::3], ref) > 1e-3


def test_displacement_unitary_block():
    """Rows n < 100 of D(4 + 4i) at cutoff 1000 are orthonormal: the truncated
    rows hold all of their weight, so this pins the large-index entries."""
    from oracle.displacement import displacement

    D = displacement((1000, 1000), 4.0 + 4.0j)
    P = D[:100] @ D[:100].conj().T
    assert np.max(np.abs(P - np.eye(100))) < 1e-12
