"""Pairwise contraction of Fock-basis arrays on the FP64 matrix cores.

Mirrors ArrayAnsatz.contract (array_ansatz.py:159-225), whose own tests
(test_array_ansatz.py) check it against einsum on the same arrays: here the
GPU result is compared with numpy's einsum of the same (sliced) arrays.  The
matrix instruction sums K in blocks of four, so the bar is normwise 1e-13.
"""

from __future__ import annotations

import numpy as np
import pytest

from tests.conftest import gpu_available

TOL = 1e-13


def rnd(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def relerr(got, ref):
    ref = np.asarray(ref)
    return np.max(np.abs(np.asarray(got) - ref)) / max(np.max(np.abs(ref)), 1e-300)


EINSUM_CASES = [
    ("ij,jk->ik", (37, 23), (23, 41)),
    ("abc,cd->abd", (4, 5, 6), (6, 7)),
    ("abcd,cdef->abef", (5, 6, 5, 6), (5, 6, 5, 6)),  # a two-mode density matrix times a two-mode operator
    ("bij,bjk->bik", (3, 20, 9), (3, 9, 130)),
    ("ij,kl->ijkl", (3, 4), (5, 2)),
    ("bij,jk->bik", (4, 6, 7), (7, 3)),
    ("ij,ij->i", (9, 70), (9, 70)),
    ("abc,ab->c", (6, 7, 8), (6, 7)),
    ("ij,jk->k", (5, 6), (6, 4)),
    ("ijk,kji->", (3, 4, 5), (5, 4, 3)),
]


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
@pytest.mark.parametrize("string,sa,sb", EINSUM_CASES, ids=lambda x: str(x))
def test_einsum_vs_numpy(string, sa, sb):
    from mrmustard_amd.contraction import einsum

    a, b = rnd(sa, 1), rnd(sb, 2)
    got = einsum(string, a, b).cpu().numpy()
    ref = np.einsum(string, a, b)
    assert got.shape == ref.shape
    assert relerr(got, ref) < TOL


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
def test_contract_array_ansatz_semantics():
    """Batch labels (str), core labels (int), a shared label with different
    sizes cut to the smaller one (array_ansatz.py:207-218)."""
    from mrmustard_amd.contraction import contract

    a, b = rnd((5, 4, 3), 3), rnd((2, 6), 4)
    got = contract(a, b, ["b", 0, 1], [1, 2], ["b", 0, 2]).cpu().numpy()
    ref = np.einsum("bij,jk->bik", a[:, :, :2], b)
    assert relerr(got, ref) < TOL
    # a batch label shared by both operands and kept
    a, b = rnd((3, 8, 8), 5), rnd((3, 8, 8), 6)
    got = contract(a, b, ["x", 0, 1], ["x", 1, 2], ["x", 2, 0]).cpu().numpy()
    assert relerr(got, np.einsum("xij,xjk->xki", a, b)) < TOL


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
def test_manager_einsum_and_batched_broadcast():
    from mrmustard_amd import math
    from mrmustard_amd.contraction import zgemm_batched_hip

    a, b = rnd((6, 10), 7), rnd((10, 9), 8)
    assert relerr(math.einsum("ij,jk->ik", a, b).cpu().numpy(), a @ b) < TOL
    A, B = rnd((1, 17, 33), 9), rnd((5, 33, 21), 10)
    assert relerr(zgemm_batched_hip(A, B).cpu().numpy(), np.matmul(np.broadcast_to(A, (5, 17, 33)), B)) < TOL


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
def test_contract_errors():
    from mrmustard_amd.contraction import contract, einsum

    a, b = rnd((3, 3), 1), rnd((3, 3), 2)
    with pytest.raises(ValueError, match="expected len"):
        contract(a, b, [0], [0, 1], [1])
    with pytest.raises(ValueError, match="Output labels"):
        contract(a, b, [0, 1], [1, 2], [3])
    # a label repeated within an operand: that operand's diagonal, as numpy's einsum
    assert relerr(contract(a, b, [0, 0], [0, 1], [1]).cpu().numpy(), np.einsum("aa,ab->b", a, b)) < TOL
    with pytest.raises(ValueError, match="repeated axis"):
        contract(rnd((3, 4), 3), b, [0, 0], [0, 1], [1])
    # einsum: a diagonal and three operands
    assert relerr(einsum("ii,ij->j", a, b).cpu().numpy(), np.einsum("ii,ij->j", a, b)) < TOL
    assert relerr(einsum("ij,jk,kl->il", a, b, a).cpu().numpy(), a @ b @ a) < TOL


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
def test_contraction_gradients_match_torch_autograd():
    """contract / einsum / zgemm_hip are differentiable like the JAX backend's
    einsum: torch grads of a real loss match torch's own autograd of the same
    einsum computed on the CPU, for both operands, with a broadcast batch."""
    import torch

    from mrmustard_amd.contraction import einsum, zgemm_batched_hip, zgemm_hip

    rng = np.random.default_rng(12)

    def cplx(*shape):
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)

    a, b2, w = cplx(3, 4, 5), cplx(5, 6), cplx(3, 4, 6)
    for fn, spec in [(einsum, "abc,cd->abd")]:
        xa = torch.tensor(a, device="cuda", requires_grad=True)
        xb = torch.tensor(b2, device="cuda", requires_grad=True)
        torch.real(torch.sum(torch.tensor(w, device="cuda") * fn(spec, xa, xb))).backward()
        ra = torch.tensor(a, requires_grad=True)
        rb = torch.tensor(b2, requires_grad=True)
        torch.real(torch.sum(torch.tensor(w) * torch.einsum(spec, ra, rb))).backward()
        assert np.max(np.abs(xa.grad.cpu().numpy() - ra.grad.numpy())) < 1e-12 * np.max(np.abs(ra.grad.numpy()))
        assert np.max(np.abs(xb.grad.cpu().numpy() - rb.grad.numpy())) < 1e-12 * np.max(np.abs(rb.grad.numpy()))
    # batched product with the right operand broadcast (batch 1)
    A3, B3, W3 = cplx(4, 7, 9), cplx(1, 9, 3), cplx(4, 7, 3)
    xa = torch.tensor(A3, device="cuda", requires_grad=True)
    xb = torch.tensor(B3, device="cuda", requires_grad=True)
    torch.real(torch.sum(torch.tensor(W3, device="cuda") * zgemm_batched_hip(xa, xb))).backward()
    ra = torch.tensor(A3, requires_grad=True)
    rb = torch.tensor(B3, requires_grad=True)
    torch.real(torch.sum(torch.tensor(W3) * torch.matmul(ra, rb))).backward()
    assert np.max(np.abs(xa.grad.cpu().numpy() - ra.grad.numpy())) < 1e-12 * np.max(np.abs(ra.grad.numpy()))
    assert xb.grad.shape == (1, 9, 3)
    assert np.max(np.abs(xb.grad.cpu().numpy() - rb.grad.numpy())) < 1e-12 * np.max(np.abs(rb.grad.numpy()))
    # plain matrices
    M1, M2 = cplx(6, 8), cplx(8, 5)
    xa = torch.tensor(M1, device="cuda", requires_grad=True)
    torch.real(torch.sum(zgemm_hip(xa, torch.tensor(M2, device="cuda")))).backward()
    ra = torch.tensor(M1, requires_grad=True)
    torch.real(torch.sum(ra @ torch.tensor(M2))).backward()
    assert np.max(np.abs(xa.grad.cpu().numpy() - ra.grad.numpy())) < 1e-12 * np.max(np.abs(ra.grad.numpy()))


def test_einsum_labels():
    """numpy's einsum string rules, on the CPU: ellipsis axes right-aligned with
    labels of their own, implicit output = labels seen once, sorted."""
    from mrmustard_amd.contraction import einsum_labels

    labs, out = einsum_labels("...abc,...c->...ab", [5, 3])
    assert labs == [["...0", "...1", "a", "b", "c"], ["...0", "...1", "c"]] and out == ["...0", "...1", "a", "b"]
    assert einsum_labels("ij,jk", [2, 2]) == ([["i", "j"], ["j", "k"]], ["i", "k"])
    assert einsum_labels("ba,c", [2, 1]) == ([["b", "a"], ["c"]], ["a", "b", "c"])
    labs, out = einsum_labels("...ab,b->...a", [4, 1])
    assert labs == [["...0", "...1", "a", "b"], ["b"]] and out == ["...0", "...1", "a"]
    with pytest.raises(ValueError):
        einsum_labels("ab,bc->ac", [2])
    with pytest.raises(ValueError):
        einsum_labels("abc->a", [2])


EINSUM_CASES = [
    ("abc,cd->abd", [(3, 4, 5), (5, 6)]),
    ("ij,jk", [(4, 5), (5, 3)]),
    ("...abc,...c->...ab", [(2, 3, 4, 5, 6), (2, 3, 6)]),
    ("ii->i", [(5, 5)]),
    ("ab->", [(4, 7)]),
    ("ba", [(3, 4)]),
    ("ij,jk,kl->il", [(3, 4), (4, 5), (5, 2)]),
    ("abc,bd,ce->ade", [(2, 3, 4), (3, 5), (4, 6)]),
    ("aab,b->a", [(4, 4, 3), (3,)]),
    ("i,i->", [(7,), (7,)]),
    ("ab,cd->abcd", [(2, 3), (4, 5)]),
    ("...ab,...b->...a", [(2, 1, 3, 4), (1, 5, 4)]),
    ("ab,cd,bc->ad", [(3, 40), (50, 2), (40, 50)]),
    ("i,ij,jk,k->", [(6,), (6, 7), (7, 8), (8,)]),
    ("i,i,j->j", [(5,), (5,), (4,)]),
    (",ij->ij", [(), (3, 4)]),
    ("ij,ij,k->k", [(3, 4), (3, 4), (6,)]),
]

# size-1 axes of any label broadcast (numpy's rule), checked against numpy
BROADCAST_CASES = [
    ("ij,ij->ij", [(1, 3), (2, 3)]),
    ("ij,jk->ik", [(2, 1), (3, 4)]),
    ("ab,b->a", [(4, 1), (6,)]),
]


def _einsum_on_cpu(monkeypatch):
    """einsum with the matrix-core product replaced by torch.matmul on the CPU,
    so that its label, diagonal and broadcasting logic runs without a GPU."""
    import torch

    import mrmustard_amd.contraction as C

    monkeypatch.setattr(C, "as_device_c128", lambda x: torch.as_tensor(x).to(torch.complex128))
    monkeypatch.setattr(C, "_zgemm_batched", lambda A, B: torch.matmul(A, B))
    return C


@pytest.mark.parametrize("spec,shapes", EINSUM_CASES + BROADCAST_CASES, ids=lambda x: str(x))
def test_einsum_logic_on_cpu(monkeypatch, spec, shapes):
    C = _einsum_on_cpu(monkeypatch)
    rng = np.random.default_rng(len(spec) + 100)
    arrs = [rng.standard_normal(s) + 1j * rng.standard_normal(s) for s in shapes]
    got = C.einsum(spec, *arrs).numpy()
    ref = np.einsum(spec, *arrs)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_einsum_and_contract_errors_on_cpu(monkeypatch):
    C = _einsum_on_cpu(monkeypatch)
    with pytest.raises(ValueError, match="sizes"):
        C.einsum("ij,jk->ik", np.zeros((2, 3)), np.zeros((4, 2)))
    with pytest.raises(ValueError, match="repeated axis"):
        C.einsum("ii->i", np.zeros((1, 3)))
    a = np.arange(9.0).reshape(3, 3)
    got = C.contract(a, np.ones((3, 2)), [0, 0], [0, 1], [1]).numpy()
    assert np.allclose(got, np.einsum("aa,ab->b", a, np.ones((3, 2))))
    with pytest.raises(ValueError, match="repeat"):
        C.contract(a, a, [0, 1], [1, 2], [0, 0])


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
@pytest.mark.parametrize("spec,shapes", EINSUM_CASES, ids=lambda x: str(x))
def test_einsum_general_vs_torch(spec, shapes):
    import torch

    from mrmustard_amd.contraction import einsum

    rng = np.random.default_rng(len(spec))
    arrs = [rng.standard_normal(s) + 1j * rng.standard_normal(s) for s in shapes]
    got = einsum(spec, *[torch.tensor(a, device="cuda") for a in arrs]).cpu().numpy()
    ref = torch.einsum(spec, *[torch.tensor(a) for a in arrs]).numpy()
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
def test_einsum_three_operands_gradient_and_errors():
    import torch

    from mrmustard_amd.contraction import einsum

    rng = np.random.default_rng(4)
    arrs = [rng.standard_normal(s) + 1j * rng.standard_normal(s) for s in [(3, 4), (4, 5), (5, 2)]]
    xs = [torch.tensor(a, device="cuda", requires_grad=True) for a in arrs]
    torch.real(torch.sum(einsum("ij,jk,kl->il", *xs))).backward()
    rs = [torch.tensor(a, requires_grad=True) for a in arrs]
    torch.real(torch.sum(torch.einsum("ij,jk,kl->il", *rs))).backward()
    for x, r in zip(xs, rs):
        assert np.max(np.abs(x.grad.cpu().numpy() - r.grad.numpy())) < 1e-12 * np.max(np.abs(r.grad.numpy()))
    with pytest.raises(ValueError, match="sizes"):
        einsum("ij,jk->ik", torch.zeros((2, 3), device="cuda"), torch.zeros((4, 2), device="cuda"))


def test_greedy_path_hand_derived():
    """opt_einsum's greedy order (restated in greedy_path; opt_einsum itself is
    not installed), on cases worked out by hand from its rules: equal label
    sets first, then the candidate removing the most memory (ties to the
    earliest operands), results keeping only labels still needed, outer
    products of the smallest last."""
    from mrmustard_amd.contraction import greedy_path

    # chain a-b-c-d, sizes a=2 b=100 c=2 d=100: (ab,bc) removes 200+200-4,
    # (bc,cd) would grow to 10^4; then the result ac meets cd
    sz = {"a": 2, "b": 100, "c": 2, "d": 100}
    assert greedy_path([["a", "b"], ["b", "c"], ["c", "d"]], ["a", "d"], sz) == [(0, 1, ["a", "c"]), (2, 3, ["a", "d"])]
    # the same chain given in another order: the cheap pair (ab, bc) is still first
    steps = greedy_path([["a", "b"], ["c", "d"], ["b", "c"]], ["a", "d"], sz)
    assert {steps[0][0], steps[0][1]} == {0, 2} and steps[0][2] == ["a", "c"]
    # equal costs: the pair with the earliest later operand wins
    eq = {x: 5 for x in "abcd"}
    assert greedy_path([["a", "b"], ["b", "c"], ["c", "d"]], ["a", "d"], eq)[0] == (0, 1, ["a", "c"])
    # equal label sets are multiplied first, then summed against the rest
    assert greedy_path([["i", "j"], ["k"], ["i", "j"]], ["k"], {"i": 3, "j": 4, "k": 5}) == [(0, 2, ["i", "j"]), (1, 3, ["k"])]
    # a label held by three operands stays in the first pair's result and
    # is summed by the second
    three = greedy_path([["x", "a"], ["x", "b"], ["x", "c"]], ["a", "b", "c"], {"x": 4, "a": 2, "b": 2, "c": 2})
    assert three == [(0, 1, ["x", "a", "b"]), (2, 3, ["a", "b", "c"])]
    # no shared label: outer products, smallest first
    assert greedy_path([["a"], ["b"], ["c"]], ["a", "b", "c"], {"a": 9, "b": 2, "c": 3}) == [(1, 2, ["b", "c"]), (0, 3, ["a", "b", "c"])]
    assert greedy_path([["a", "b"]], ["a"], {"a": 2, "b": 2}) == []


def test_contract_middle_block_on_cpu(monkeypatch):
    """A wire block in the middle of the larger operand, with the output
    keeping its other axes in place (the ket side of an operator on a density
    matrix): read in place, same result as numpy."""
    C = _einsum_on_cpu(monkeypatch)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((2, 3, 4, 40)) + 1j * rng.standard_normal((2, 3, 4, 40))
    y = rng.standard_normal((4, 3, 6)) + 1j * rng.standard_normal((4, 3, 6))
    got = C.contract(x, y, [0, 1, 2, 3], [2, 1, "y"], [0, "y", 3]).numpy()
    assert np.allclose(got, np.einsum("abcd,cby->ayd", x, y), rtol=1e-13, atol=1e-13)
    assert C._middle_block(torch_t(x), [0, 1, 2, 3], torch_t(y), [2, 1, "y"], [0, "y", 3]) is not None
    # the output in another order is not that form: the generic path
    got2 = C.contract(x, y, [0, 1, 2, 3], [2, 1, "y"], [3, 0, "y"]).numpy()
    assert np.allclose(got2, np.einsum("abcd,cby->day", x, y), rtol=1e-13, atol=1e-13)
    assert C._middle_block(torch_t(x), [0, 1, 2, 3], torch_t(y), [2, 1, "y"], [3, 0, "y"]) is None
    # summed axes not adjacent in x: not that form
    z = rng.standard_normal((3, 40, 2)) + 1j * rng.standard_normal((3, 40, 2))
    # trailing summed wires: the generic product, which reads x in place
    w = rng.standard_normal((4, 40, 7)) + 1j * rng.standard_normal((4, 40, 7))
    assert C._middle_block(torch_t(x), [0, 1, 2, 3], torch_t(w), [2, 3, "q"], [0, 1, "q"]) is None
    assert C._middle_block(torch_t(x), [0, 1, 2, 3], torch_t(z), [1, 3, "z"], [0, "z", 2]) is None


def torch_t(a):
    import torch

    return torch.as_tensor(a)
