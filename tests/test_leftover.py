"""The one-leftover-mode strategy: the oracle pinned on the CPU, the HIP path
against it on the GPU.

Mirrors test_compactFock.py:46-69 and test_fast_diagonal.py:24-45: the
conditional density matrices out[m, n, k] equal the entries G[(m, k), (n, k)]
of the full lattice of the same triple.  A product state pins the rest in
closed form: a coherent mode 0 beside thermal detected modes gives
out[m, n, k] = rho_coherent[m, n] * prod_j P_thermal(k_j).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from tests.conftest import gpu_available
from tests.test_diagonal import random_dm

TOL = 1e-12  # normwise: the inner products' summation order differs from the reference's BLAS dot


def full_leftover(full, o, pnr):
    """out[m, n, k] = full[(m, k), (n, k)] with full indexed [mode0, tail..., mode0, tail...]."""
    out = np.zeros((o + 1, o + 1, *(k + 1 for k in pnr)), complex)
    for k in np.ndindex(*(p + 1 for p in pnr)):
        out[(slice(None), slice(None)) + k] = full[(slice(None),) + k + (slice(None),) + k]
    return out


def product_state(alpha, nbars):
    """Coherent mode 0 and thermal modes 1..: (A, b, c) in [bra modes..., ket modes...]."""
    M = 1 + len(nbars)
    A = np.zeros((2 * M, 2 * M), complex)
    for j, n in enumerate(nbars, start=1):
        A[j, M + j] = A[M + j, j] = n / (1 + n)
    b = np.zeros(2 * M, complex)
    b[0], b[M] = np.conj(alpha), alpha
    c = math.exp(-abs(alpha) ** 2) / np.prod([1 + n for n in nbars])
    return A, b, c


def product_closed_form(alpha, nbars, o, pnr):
    m = np.arange(o + 1)
    fact = np.array([math.sqrt(math.factorial(k)) for k in m])
    rho = math.exp(-abs(alpha) ** 2) * np.outer(np.conj(alpha) ** m / fact, alpha**m / fact)
    out = rho.reshape(o + 1, o + 1, *([1] * len(pnr)))
    for j, (n, k) in enumerate(zip(nbars, pnr)):
        pk = (n / (1 + n)) ** np.arange(k + 1) / (1 + n)
        shape = [1] * (2 + len(pnr))
        shape[2 + j] = k + 1
        out = out * pk.reshape(shape)
    return out


def relerr(got, ref):
    return np.max(np.abs(np.asarray(got) - ref)) / max(np.max(np.abs(ref)), 1e-300)


CPU_CASES = [(2, 3, (4,), 1), (3, 3, (1, 2), 2), (3, 2, (3, 2), 3), (4, 2, (2, 1, 2), 4), (2, 0, (3,), 5), (3, 3, (0, 2), 6)]


# ----------------------------------------------------------------- oracle


@pytest.mark.parametrize("M,o,pnr,seed", CPU_CASES)
def test_oracle_vs_full_lattice(oracle_mod, M, o, pnr, seed):
    """test_compactFock.py:46-69 with the pinned vanilla oracle as the full lattice."""
    from oracle import leftover as OL

    A, b, c = random_dm(M, seed)
    shape = (o + 1,) + tuple(k + 1 for k in pnr)
    ref = full_leftover(oracle_mod.vanilla(shape * 2, A, b, c), o, pnr)
    assert relerr(OL.leftover(A, b, c, o, pnr), ref) < 1e-13


def test_oracle_product_state():
    from oracle import leftover as OL

    alpha, nbars, o, pnr = 0.8 - 0.5j, (0.7, 1.5), 6, (5, 4)
    got = OL.leftover(*product_state(alpha, nbars), o, pnr)
    assert np.max(np.abs(got - product_closed_form(alpha, nbars, o, pnr))) < 1e-15


# ----------------------------------------------------------------- GPU

GPU_CASES = CPU_CASES + [(2, 5, (10,), 7), (3, 3, (4, 5), 8), (5, 2, (2, 2, 2, 2), 9), (8, 1, (1,) * 7, 10)]


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
@pytest.mark.parametrize("M,o,pnr,seed", GPU_CASES, ids=lambda x: str(x))
def test_leftover_hip_vs_oracle(M, o, pnr, seed):
    from oracle import diagonal as OD
    from oracle import leftover as OL

    from mrmustard_amd.lattice.compact_fock import leftover_hip

    A, b, c = random_dm(M, seed)
    print(f"1leftover M={M} output_cutoff={o} pnr={pnr}", flush=True)
    ref = OL.leftover(A, b, c, o, pnr)
    got = leftover_hip(A, b, c, o, pnr, reordered=True).cpu().numpy()
    assert got.shape == ref.shape
    assert relerr(got, ref) < TOL
    Ar, br = OD.reorder_ab(A, b)
    got2 = leftover_hip(Ar, br, c, o, pnr, reordered=False).cpu().numpy()
    assert relerr(got2, ref) < TOL


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
def test_leftover_hip_product_state():
    from mrmustard_amd.lattice.compact_fock import leftover_hip

    alpha, nbars, o, pnr = 1.3 + 0.4j, (0.5, 2.0), 20, (30, 25)
    got = leftover_hip(*product_state(alpha, nbars), o, pnr).cpu().numpy()
    assert np.max(np.abs(got - product_closed_form(alpha, nbars, o, pnr))) < 1e-14
    # a leftover block wider than one wave (101 x 101)
    alpha, nbars, o, pnr = 2.5 - 1.0j, (1.0,), 100, (3,)
    got = leftover_hip(*product_state(alpha, nbars), o, pnr).cpu().numpy()
    assert np.max(np.abs(got - product_closed_form(alpha, nbars, o, pnr))) < 1e-14


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
@pytest.mark.parametrize("M,o,pnr", [(2, 10, (20,)), (3, 4, (8, 8))])
def test_leftover_hip_vs_full_lattice_hip(M, o, pnr):
    """test_fast_diagonal.py:24-45 at sizes beyond the oracle: the full lattice
    from the (bit-exact) vanilla kernel, through the manager."""
    from mrmustard_amd import math as mm

    A, b, c = random_dm(M, 60 + M)
    shape = (o + 1,) + tuple(k + 1 for k in pnr)
    ref = full_leftover(mm.hermite_renormalized(A, b, c, shape=shape * 2).cpu().numpy(), o, pnr)
    got = mm.hermite_renormalized_1leftoverMode(A, b, c, o, pnr).cpu().numpy()
    assert relerr(got, ref) < TOL
    # a density matrix per detection pattern: Hermitian blocks
    assert np.max(np.abs(got - np.conj(np.swapaxes(got, 0, 1)))) < 1e-12 * np.max(np.abs(got))


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
def test_leftover_hip_errors():
    from mrmustard_amd.lattice.compact_fock import leftover_hip

    A, b, c = random_dm(2, 1)
    with pytest.raises(ValueError, match="incompatible"):
        leftover_hip(A, b, c, 3, (2, 2))
    bad = np.array(A)
    bad[0, 1] += 1.0
    with pytest.raises(ValueError, match="symmetric"):
        leftover_hip(bad, b, c, 3, (2,))
    A1, b1, c1 = random_dm(1, 1)
    with pytest.raises(ValueError, match="greater than 1"):
        leftover_hip(A1, b1, c1, 3, ())
    A9, b9, c9 = product_state(0.1, (0.1,) * 8)
    with pytest.raises(ValueError):
        leftover_hip(A9, b9, c9, 1, (1,) * 8)


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
@pytest.mark.parametrize("M,o,pnr", [(2, 3, (0,)), (3, 2, (0, 0)), (3, 0, (0, 3)), (2, 0, (0,))], ids=lambda x: str(x))
def test_leftover_hip_trivial_boxes(M, o, pnr):
    """Zero detected photons allowed (PNR cutoffs 0) and a one-entry leftover block."""
    from oracle import leftover as OL

    from mrmustard_amd.lattice.compact_fock import leftover_hip

    A, b, c = random_dm(M, 95 + M)
    ref = OL.leftover(A, b, c, o, pnr)
    got = leftover_hip(A, b, c, o, pnr).cpu().numpy()
    assert got.shape == ref.shape and relerr(got, ref) < TOL


# ----------------------------------------------------------------- VJP

VJP_CASES = [(2, 3, (4,), 11), (3, 3, (1, 2), 12), (3, 2, (3, 2), 13), (4, 2, (2, 1, 2), 14), (2, 0, (3,), 15), (3, 3, (0, 2), 16)]


def caller_order(dA_c, dB_c, M):
    """Compact-order gradients mapped to the [modes..., modes...] order (inverse of reorder_ab)."""
    order = np.arange(2 * M).reshape(2, -1).T.flatten()
    dA = np.zeros_like(dA_c)
    dA[np.ix_(order, order)] = dA_c
    dB = np.zeros_like(dB_c)
    dB[order] = dB_c
    return dA, dB


@pytest.mark.parametrize("M,o,pnr,seed", VJP_CASES[:4])
def test_oracle_vjp_vs_finite_differences(M, o, pnr, seed):
    """The forward-mode Jacobian (singleLeftoverMode_grad.py) against central
    differences of the pinned forward, every entry of A and b perturbed on its
    own, and c; a dropped or misplaced derivative term fails here."""
    from oracle import diagonal as OD
    from oracle import leftover as OL

    A, b, c = random_dm(M, seed)
    Ar, br = OD.reorder_ab(A, b)
    cut = (o + 1,) + tuple(k + 1 for k in pnr)
    val, _, _ = OL.leftover_jacobian(Ar, br, c, cut)
    assert relerr(val, OL.leftover_compact(Ar, br, c, cut)) < 1e-14
    rng = np.random.default_rng(seed)
    cot = rng.standard_normal(val.shape) + 1j * rng.standard_normal(val.shape)
    dA, dB, dc = OL.leftover_vjp(Ar, br, c, cut, cot)

    def loss(A_, B_, c_):
        return np.sum(cot * OL.leftover_compact(A_, B_, c_, cut))

    h, G = 1e-6, 2 * M
    fdA = np.zeros((G, G), complex)
    for i in range(G):
        for j in range(G):
            E = np.zeros((G, G), complex)
            E[i, j] = h
            fdA[i, j] = (loss(Ar + E, br, c) - loss(Ar - E, br, c)) / (2 * h)
    fdB = np.array([(loss(Ar, br + h * np.eye(G)[i], c) - loss(Ar, br - h * np.eye(G)[i], c)) / (2 * h) for i in range(G)])
    fdc = (loss(Ar, br, c + h) - loss(Ar, br, c - h)) / (2 * h)
    assert relerr(dA, fdA) < 1e-8 and relerr(dB, fdB) < 1e-8 and abs(dc - fdc) < 1e-8 * abs(dc)


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
@pytest.mark.parametrize("M,o,pnr,seed", VJP_CASES + [(2, 6, (9,), 17), (3, 4, (5, 4), 18), (5, 1, (2, 1, 2, 1), 19)],
                         ids=lambda x: str(x))
def test_leftover_vjp_hip_vs_oracle(M, o, pnr, seed):
    from oracle import diagonal as OD
    from oracle import leftover as OL

    from mrmustard_amd.lattice.compact_fock import leftover_vjp_hip

    A, b, c = random_dm(M, seed)
    Ar, br = OD.reorder_ab(A, b)
    cut = (o + 1,) + tuple(k + 1 for k in pnr)
    rng = np.random.default_rng(seed)
    shape = (o + 1,) + cut
    cot = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    print(f"1leftover VJP M={M} output_cutoff={o} pnr={pnr}", flush=True)
    dA_c, dB_c, dc = OL.leftover_vjp(Ar, br, c, cut, cot)
    out, gA, gB, gc = leftover_vjp_hip(Ar, br, c, o, pnr, cot, reordered=False)
    assert relerr(out.cpu().numpy(), OL.leftover_compact(Ar, br, c, cut)) < TOL
    assert relerr(gA.cpu().numpy(), dA_c) < TOL and relerr(gB.cpu().numpy(), dB_c) < TOL
    assert abs(complex(gc) - dc) <= TOL * abs(dc)
    dA, dB = caller_order(dA_c, dB_c, M)
    _, gA, gB, gc = leftover_vjp_hip(A, b, c, o, pnr, cot, reordered=True)
    assert relerr(gA.cpu().numpy(), dA) < TOL and relerr(gB.cpu().numpy(), dB) < TOL
    # deterministic: a second call gives the same bits
    _, gA2, gB2, _ = leftover_vjp_hip(A, b, c, o, pnr, cot, reordered=True)
    assert np.array_equal(gA.cpu().numpy(), gA2.cpu().numpy()) and np.array_equal(gB.cpu().numpy(), gB2.cpu().numpy())


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
def test_leftover_autograd_convention():
    """torch grads = conj(VJP_ref(conj(grad_out))) through
    math.hermite_renormalized_1leftoverMode (the reference's JAX custom VJP is
    holomorphic)."""
    import torch

    from oracle import diagonal as OD
    from oracle import leftover as OL

    from mrmustard_amd import math as mm

    M, o, pnr = 3, 3, (2, 3)
    A, b, c = random_dm(M, 21)
    At = torch.tensor(np.asarray(A), device="cuda", requires_grad=True)
    bt = torch.tensor(np.asarray(b), device="cuda", requires_grad=True)
    ct = torch.tensor(complex(c), dtype=torch.complex128, device="cuda", requires_grad=True)
    out = mm.hermite_renormalized_1leftoverMode(At, bt, ct, o, pnr)
    rng = np.random.default_rng(5)
    w = rng.standard_normal(tuple(out.shape)) + 1j * rng.standard_normal(tuple(out.shape))
    torch.real(torch.sum(torch.tensor(w, device="cuda") * out)).backward()
    Ar, br = OD.reorder_ab(A, b)
    dA_c, dB_c, dc = OL.leftover_vjp(Ar, br, c, (o + 1,) + tuple(k + 1 for k in pnr), w)
    dA, dB = caller_order(dA_c, dB_c, M)
    assert relerr(At.grad.cpu().numpy(), np.conj(dA)) < TOL
    assert relerr(bt.grad.cpu().numpy(), np.conj(dB)) < TOL
    assert abs(complex(ct.grad) - np.conj(dc)) <= TOL * abs(dc)


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
@pytest.mark.parametrize("M,o,pnr", [(3, 20, (14, 14)), (2, 100, (4,))], ids=lambda x: str(x))
def test_leftover_vjp_large_vs_finite_differences(M, o, pnr):
    """Sizes beyond the oracle (the bench's 21 x 21 x 15 x 15, a 101 x 101
    block): the VJP of the real loss sum(w * out) in random directions of
    (A, b, c) against central differences of the GPU forward, A perturbed
    symmetrically (a physical direction)."""
    from mrmustard_amd.lattice.compact_fock import leftover_hip, leftover_vjp_hip

    A, b, c = random_dm(M, 40 + M)
    shape = (o + 1, o + 1, *(k + 1 for k in pnr))
    rng = np.random.default_rng(6)
    w = rng.standard_normal(shape)
    _, gA, gB, gc = leftover_vjp_hip(A, b, c, o, pnr, w.astype(complex))
    gA, gB, gc = gA.cpu().numpy(), gB.cpu().numpy(), complex(gc)

    def loss(A_, b_, c_):
        return np.sum(w * leftover_hip(A_, b_, c_, o, pnr).cpu().numpy())

    h, G = 1e-6, 2 * M
    for k in range(3):
        dA = rng.standard_normal((G, G)) + 1j * rng.standard_normal((G, G))
        dA = (dA + dA.T) / 2
        db = rng.standard_normal(G) + 1j * rng.standard_normal(G)
        dc = complex(rng.standard_normal(), rng.standard_normal())
        fd = (loss(A + h * dA, b + h * db, c + h * dc) - loss(A - h * dA, b - h * db, c - h * dc)) / (2 * h)
        an = np.sum(gA * dA) + np.sum(gB * db) + gc * dc
        assert abs(an - fd) < 1e-6 * max(1.0, abs(fd)), (k, an, fd)
