"""Seeded synthetic (A, b, c) workloads for the Bargmann->Fock path.

Each function builds complex128 numpy inputs of the kind the reference's
state constructors produce:

- ``gaussian_ket``: A = U diag(tanh r) U^T with U Haar-random, b ~ CN(0, sigma^2),
  c = 1.  A pure Gaussian ket triple.
- ``lossy_dm``: the ket of ``gaussian_ket`` as a density matrix, sent through
  per-mode pure loss, as a 2m-index triple with the bra indices first
  (index order of physics/wires.py:101-106, bra wires before ket wires).
- ``squeezed_interferometer_dm``: single-mode squeezed vacua (triples.py:126-150)
  through a Haar interferometer, displaced (the displacement rule of
  triples.py:153-190 applied to m modes) and then lossy.

The loss map is derived, not copied.  With F(w, z) the generating function of
the density matrix (bra variables w, ket variables z), per-mode pure loss of
transmissivity eta acts as
    F_out(w, z) = [exp((1 - eta) d_w d_z) F](sqrt(eta) w, sqrt(eta) z),
(apply the Kraus operators sqrt((1-eta)^l / l!) eta^{n/2} a^l to Fock
matrix elements and resum).  On a Gaussian exp(v^T M v / 2 + beta^T v) the
heat operator exp(d^T T d / 2) gives
    M' = M (I - T M)^-1,  beta' = (I - M T)^-1 beta,
    c' = c det(I - T M)^-1/2 exp(beta^T T (I - M T)^-1 beta / 2).
The test suite checks the result against the Kraus sum in the Fock basis.
"""

from __future__ import annotations

import numpy as np


def haar_unitary(m: int, rng: np.random.Generator) -> np.ndarray:
    """Haar-random m x m unitary: QR of a complex Ginibre matrix, phases fixed."""
    z = (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r)
    return q * (d / np.abs(d))


def complex_normal(shape, sigma: float, rng: np.random.Generator) -> np.ndarray:
    """CN(0, sigma^2): real and imaginary parts each N(0, sigma^2 / 2)."""
    s = sigma / np.sqrt(2.0)
    return rng.normal(0.0, s, shape) + 1j * rng.normal(0.0, s, shape)


def gaussian_ket(m: int, rng: np.random.Generator, r_max: float = 1.0, b_sigma: float = 0.5):
    """Pure Gaussian ket triple on m modes (A = U diag(tanh r) U^T, c = 1)."""
    U = haar_unitary(m, rng)
    r = rng.uniform(0.0, r_max, m)
    A = U @ np.diag(np.tanh(r)) @ U.T
    b = complex_normal(m, b_sigma, rng)
    return A.astype(np.complex128), b.astype(np.complex128), np.complex128(1.0)


def ket_to_dm(A, b, c):
    """|psi><psi| as a triple with bra indices first: (conj A (+) A, [conj b, b], |c|^2)."""
    m = A.shape[0]
    Z = np.zeros((m, m), np.complex128)
    Adm = np.block([[np.conj(A), Z], [Z, A]])
    bdm = np.concatenate([np.conj(b), b])
    return Adm, bdm, np.complex128(abs(complex(c)) ** 2)


def apply_loss(A, b, c, etas):
    """Per-mode pure loss on a DM triple with index order [bra_0..bra_{m-1}, ket_0..ket_{m-1}]."""
    A = np.asarray(A, np.complex128)
    b = np.asarray(b, np.complex128)
    etas = np.asarray(etas, np.float64)
    m = len(etas)
    n = 2 * m
    T = np.zeros((n, n), np.complex128)
    for i, eta in enumerate(etas):
        T[i, m + i] = T[m + i, i] = 1.0 - eta
    I = np.eye(n)
    ITM = I - T @ A
    IMT = I - A @ T
    A1 = A @ np.linalg.inv(ITM)
    b1 = np.linalg.solve(IMT, b)
    c1 = complex(c) / np.sqrt(np.linalg.det(ITM).astype(np.complex128)) * np.exp(0.5 * b @ T @ b1)
    scale = np.sqrt(np.concatenate([etas, etas]))
    A2 = (A1 * scale[:, None]) * scale[None, :]
    b2 = b1 * scale
    A2 = 0.5 * (A2 + A2.T)
    return A2, b2, np.complex128(c1)


def displace_ket(A, b, c, alpha):
    """Displacement D(alpha) applied to a ket triple: the Bargmann function
    f(z) -> exp(-|alpha|^2/2 + alpha^T z) f(z - conj(alpha))."""
    alpha = np.asarray(alpha, np.complex128)
    ac = np.conj(alpha)
    b2 = b + alpha - A @ ac
    c2 = complex(c) * np.exp(-0.5 * np.sum(np.abs(alpha) ** 2) + 0.5 * ac @ A @ ac - b @ ac)
    return A, b2, np.complex128(c2)


# -------------------------------------------------------------- configurations


def plain_ket_50x50(seed: int = 0):
    """2-mode Gaussian ket, shape (50, 50): 2,500 amplitudes."""
    rng = np.random.default_rng(seed)
    A, b, c = gaussian_ket(2, rng)
    return (A, b, c), (50, 50)


def lossy_dm_40x4(seed: int = 1):
    """2-mode Gaussian ket through per-mode loss eta ~ U[0.7, 1], as a 4-index
    density-matrix triple, shape (40, 40, 40, 40): 2.56 M amplitudes."""
    rng = np.random.default_rng(seed)
    A, b, c = gaussian_ket(2, rng)
    etas = rng.uniform(0.7, 1.0, 2)
    return apply_loss(*ket_to_dm(A, b, c), etas), (40, 40, 40, 40)


def uneven_kets(n: int = 4096, seed: int = 2, max_amps: int = 200_000, cut_lo: int = 8, cut_hi: int = 60):
    """``n`` Gaussian kets with m ~ U{1,2,3} modes and per-mode cutoffs ~
    U{cut_lo..cut_hi}; while the product exceeds ``max_amps`` the largest
    cutoff (first one on ties) is decremented."""
    rng = np.random.default_rng(seed)
    triples, shapes = [], []
    for _ in range(n):
        m = int(rng.integers(1, 4))
        triples.append(gaussian_ket(m, rng))
        cut = [int(x) for x in rng.integers(cut_lo, cut_hi + 1, m)]
        while int(np.prod(cut)) > max_amps:
            cut[int(np.argmax(cut))] -= 1
        shapes.append(tuple(cut))
    return triples, shapes


def batched_kets_30x3(n: int = 256, seed: int = 3, with_cotangent: bool = True):
    """``n`` three-mode kets (r ~ U[0, 1.2], b ~ CN(0, 0.5^2)), shape (30, 30, 30)
    each, plus an upstream cotangent dL/dG ~ CN(0, 1) per amplitude."""
    rng = np.random.default_rng(seed)
    A = np.empty((n, 3, 3), np.complex128)
    b = np.empty((n, 3), np.complex128)
    c = np.ones(n, np.complex128)
    for k in range(n):
        A[k], b[k], c[k] = gaussian_ket(3, rng, r_max=1.2)
    shape = (30, 30, 30)
    g = complex_normal((n,) + shape, 1.0, rng) if with_cotangent else None
    return (A, b, c), shape, g


def squeezed_interferometer_dm(m: int = 4, seed: int = 4):
    """m single-mode squeezed vacua (r ~ U[0.2, 0.8]) through a Haar
    interferometer, displaced by alpha ~ CN(0, 0.3^2), then loss eta ~ U[0.8, 1]:
    a 2m-index density-matrix triple."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(0.2, 0.8, m)
    A0 = np.diag(-np.tanh(r)).astype(np.complex128)
    c0 = np.complex128(np.prod(1.0 / np.sqrt(np.cosh(r))))
    U = haar_unitary(m, rng)
    A = U @ A0 @ U.T
    b = np.zeros(m, np.complex128)
    alpha = complex_normal(m, 0.3, rng)
    A, b, c = displace_ket(A, b, c0, alpha)
    etas = rng.uniform(0.8, 1.0, m)
    return apply_loss(*ket_to_dm(A, b, c), etas)


def dm8_10x8(seed: int = 4):
    """4-mode lossy squeezed-interferometer density matrix, shape (10,)*8: 1e8 amplitudes."""
    return squeezed_interferometer_dm(4, seed), (10,) * 8


def normalize_dm(A, b, c):
    """Scale c so that the density matrix has unit trace:
    Tr = c exp(-b^T (A - X)^-1 b / 2) / sqrt(det(A - X)), X = [[0, I], [I, 0]]
    (the Gaussian integral over all modes, autoshape.py:105-108 applied to all of them)."""
    A = np.asarray(A, np.complex128)
    b = np.asarray(b, np.complex128)
    M = len(b) // 2
    X = np.block([[np.zeros((M, M)), np.eye(M)], [np.eye(M), np.zeros((M, M))]])
    R = A - X
    tr = complex(c) * np.exp(-0.5 * b @ np.linalg.solve(R, b)) / np.sqrt(np.linalg.det(R))
    return A, b, np.complex128(complex(c) / abs(tr))


def autoshape_states(n: int = 4096, m: int = 4, seed: int = 5):
    """n normalized lossy Gaussian density matrices of m modes (A: (n, 2m, 2m),
    b: (n, 2m), c: (n,)), the input of a batched autoshape."""
    rng = np.random.default_rng(seed)
    As, bs, cs = [], [], []
    for _ in range(n):
        A, b, c = gaussian_ket(m, rng, r_max=0.8, b_sigma=0.6)
        A, b, c = normalize_dm(*apply_loss(*ket_to_dm(A, b, c), rng.uniform(0.6, 1.0, m)))
        As.append(A)
        bs.append(b)
        cs.append(c)
    return np.stack(As), np.stack(bs), np.array(cs)


def pnr_dm(m: int = 6, seed: int = 6):
    """A normalised m-mode lossy Gaussian density matrix (A, b, c) in the
    [bra modes..., ket modes...] order: its diagonal lattice entries are the
    joint photon-number probabilities."""
    rng = np.random.default_rng(seed)
    A, b, c = gaussian_ket(m, rng, r_max=0.6, b_sigma=0.5)
    dm = apply_loss(*ket_to_dm(A, b, c), rng.uniform(0.6, 1.0, m))
    return normalize_dm(*dm)


def displacement_gate_triple(alpha: complex):
    """(A, b, c) of the one-mode displacement gate D(alpha) on wires (out, in)
    (triples.py:424-446): A = [[0, 1], [1, 0]], b = [alpha, -conj(alpha)],
    c = exp(-|alpha|^2 / 2)."""
    alpha = complex(alpha)
    A = np.array([[0.0, 1.0], [1.0, 0.0]], np.complex128)
    b = np.array([alpha, -np.conj(alpha)], np.complex128)
    return A, b, np.complex128(np.exp(-abs(alpha) ** 2 / 2))


def squeezing_gate_triple(r: float, phi: float = 0.0):
    """(A, b, c) of the one-mode squeezing gate S(r, phi) on wires (out, in)
    (triples.py:449-483)."""
    t, s = np.tanh(r), 1.0 / np.cosh(r)
    A = np.array([[-np.exp(1j * phi) * t, s], [s, np.exp(-1j * phi) * t]], np.complex128)
    return A, np.zeros(2, np.complex128), np.complex128(1.0 / np.sqrt(np.cosh(r)))
