"""A Bargmann triple converted to Fock and contracted with a Fock array in one
pass (CircuitComponent.contract's to_fock + ArrayAnsatz.contract,
circuit_components.py:445-458, array_ansatz.py:159-225).

Oracle: the pinned vanilla oracle's lattice contracted by numpy's einsum.  The
fused pass reorders the triple's variables and sums in its own order, so the
bar is 1e-12 normwise; it must also equal the unfused path (whole lattice,
then ``contract``) to that bar.
"""

from __future__ import annotations

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

TOL = 1e-12


def relerr(got, ref):
    got, ref = np.asarray(got), np.asarray(ref)
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


def triple(D, seed, batch=()):
    rng = np.random.default_rng(seed)

    def one():
        X = rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D))
        A = 0.35 * (X + X.T) / np.sqrt(D)
        b = 0.3 * (rng.standard_normal(D) + 1j * rng.standard_normal(D))
        c = complex(rng.standard_normal(), rng.standard_normal())
        return A, b, c

    if not batch:
        return one()
    parts = [one() for _ in range(int(np.prod(batch)))]
    return (np.stack([p[0] for p in parts]).reshape(batch + (D, D)), np.stack([p[1] for p in parts]).reshape(batch + (D,)),
            np.array([p[2] for p in parts]).reshape(batch))


def oracle_lattice(oracle_mod, A, b, c, shape):
    A, b, c = np.asarray(A), np.asarray(b), np.asarray(c)
    if A.ndim == 2:
        return oracle_mod.vanilla(tuple(shape), A, b, complex(c))
    B = A.shape[:-2]
    flat = [oracle_mod.vanilla(tuple(shape), a, bb, complex(cc)) for a, bb, cc in
            zip(A.reshape(-1, *A.shape[-2:]), b.reshape(-1, b.shape[-1]), c.reshape(-1))]
    return np.stack(flat).reshape(B + tuple(shape))


def letters(labels):
    table = {}
    out = []
    for lab in labels:
        if lab not in table:
            table[lab] = "abcdefghijklmnopqrstuvwxyz"[len(table)]
        out.append(table[lab])
    return out, table


def np_contract(G, W, idx1, idx2, idx_out):
    shared = set(idx1) & set(idx2)
    sl1 = [slice(None)] * G.ndim
    sl2 = [slice(None)] * W.ndim
    for lab in shared:
        d = min(G.shape[idx1.index(lab)], W.shape[idx2.index(lab)])
        sl1[idx1.index(lab)] = slice(0, d)
        sl2[idx2.index(lab)] = slice(0, d)
    G, W = G[tuple(sl1)], W[tuple(sl2)]
    _, table = letters(list(idx1) + list(idx2))
    s = "".join(table[x] for x in idx1) + "," + "".join(table[x] for x in idx2) + "->" + "".join(table[x] for x in idx_out)
    return np.einsum(s, G, W)


CASES = [
    # (D, shape, batch, idx1, array shape, idx2, idx_out)
    (3, (20, 20, 20), (), [0, 1, 2], (20, 20, 5), [1, 2, "x"], [0, "x"]),            # trailing wires, M = 5
    (3, (12, 15, 18), (), [0, 1, 2], (15, 7), [1, "x"], [0, 2, "x"]),                # a middle wire: variables reordered
    (3, (10, 11, 12), (), [0, 1, 2], (12,), [2], [1, 0]),                            # a vector, output transposed
    (2, (30, 40), (3,), ["b", 0, 1], (25, 6), [1, "y"], ["b", 0, "y"]),              # a batch, the array cut to 25
    (3, (4, 80, 80), (), [0, 1, 2], (80, 80), [2, "z"], [0, 1, "z"]),                # slabs beyond LDS: window + matrix cores
    (4, (3, 40, 45, 40), (), [0, 1, 2, 3], (45, 40, 4), [2, 3, "w"], ["w", 1, 0]),   # window, two wires contracted
    (4, (6, 6, 6, 6), (), [0, 1, 2, 3], (6, 6, 6, 6), [2, 3, "p", "q"], [0, 1, "p", "q"]),  # an operator on two wires
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[1])))
def test_fock_contract_vs_oracle(oracle_mod, case):
    from mrmustard_amd.lattice.fused import fock_contract

    D, shape, batch, idx1, ashape, idx2, idx_out = case
    A, b, c = triple(D, sum(shape), batch)
    rng = np.random.default_rng(len(ashape))
    W = rng.standard_normal(ashape) + 1j * rng.standard_normal(ashape)
    print(f"fock_contract shape={shape} batch={batch} array={ashape}", flush=True)
    ref = np_contract(oracle_lattice(oracle_mod, A, b, c, shape), W, idx1, idx2, idx_out)
    got = fock_contract(A, b, c, shape, idx1, W, idx2, idx_out).cpu().numpy()
    assert got.shape == ref.shape
    assert relerr(got, ref) < TOL
    unfused = fock_contract(A, b, c, shape, idx1, W, idx2, idx_out, fused=False).cpu().numpy()
    assert relerr(got, unfused) < TOL


def test_fock_contract_uncovered_cases_fall_back(oracle_mod):
    """A label shared and kept (a zip), a wire summed on the lattice side only,
    and tracked gradients take the unfused path with the same results."""
    import torch

    from mrmustard_amd.lattice.fused import fock_contract

    A, b, c = triple(3, 5)
    shape = (6, 7, 8)
    G = oracle_lattice(oracle_mod, A, b, c, shape)
    rng = np.random.default_rng(9)
    W = rng.standard_normal((8, 3)) + 1j * rng.standard_normal((8, 3))
    for idx2, idx_out in (([2, "x"], [0, 1, 2, "x"]), ([2, "x"], [0, "x"])):
        got = fock_contract(A, b, c, shape, [0, 1, 2], W, idx2, idx_out).cpu().numpy()
        assert relerr(got, np_contract(G, W, [0, 1, 2], idx2, idx_out)) < TOL
    At = torch.tensor(A, device="cuda", requires_grad=True)
    out = fock_contract(At, b, c, shape, [0, 1, 2], W, [2, "x"], [0, 1, "x"])
    out.abs().sum().backward()
    assert At.grad is not None and torch.isfinite(At.grad).all()
