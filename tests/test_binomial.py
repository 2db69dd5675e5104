"""The binomial strategy: oracle properties on the CPU, the HIP path against it on the GPU.

Mirrors test_lattice_functions.py:33-45 (binomial vs its dict form: the same
values) and :122-138 (binomial equals vanilla on the filled part).
"""

from __future__ import annotations

import numpy as np
import pytest

from mrmustard_amd import workloads as W
from tests.conftest import gpu_available


def normalized_ket(m, seed):
    rng = np.random.default_rng(seed)
    A, b, _ = W.gaussian_ket(m, rng, r_max=0.7, b_sigma=0.5)
    _, _, c_dm = W.normalize_dm(*W.ket_to_dm(A, b, 1.0))
    return A, b, np.sqrt(abs(complex(c_dm)))


def stop_weight(G, shape):
    w = np.sum(np.indices(shape), axis=0)
    nz = w[np.abs(G) > 0]
    return int(nz.max()) if nz.size else 0


# ----------------------------------------------------------------- oracle


def test_oracle_no_stop_equals_vanilla(oracle_mod):
    from oracle import binomial as OB

    A, b, c = normalized_ket(3, 1)
    shape = (6, 5, 7)
    G, norm = OB.binomial(shape, A, b, c, None, sum(shape) - len(shape) + 1)
    full = oracle_mod.vanilla(shape, A, b, c)
    assert np.array_equal(G, full)
    assert abs(norm - np.sum(np.abs(full) ** 2)) < 1e-12


def test_oracle_vanilla_vs_binomial_reference_case(oracle_mod):
    """test_lattice_functions.py:122-138: shapes (5,5), max_l2 0.9999, global cutoff 12."""
    from oracle import binomial as OB

    A, b, c = normalized_ket(2, 42)
    G, _ = OB.binomial((5, 5), A, b, c, 0.9999, 12)
    w = stop_weight(G, (5, 5))
    full = oracle_mod.vanilla((10, 10), A, b, c)[:5, :5]
    keep = np.sum(np.indices((5, 5)), axis=0) <= w
    assert np.array_equal(G[keep], full[keep]) and not np.any(G[~keep])


@pytest.mark.parametrize("max_l2", [0.5, 0.8, 0.9, 0.99])
def test_oracle_stop_rule(oracle_mod, max_l2):
    """The fill stops at the first weight whose running squared norm passes max_l2."""
    from oracle import binomial as OB

    A, b, c = normalized_ket(2, 7)
    shape = (30, 30)
    G, norm = OB.binomial(shape, A, b, c, max_l2, 59)
    full = oracle_mod.vanilla(shape, A, b, c)
    weights = np.sum(np.indices(shape), axis=0)
    w = stop_weight(G, shape)
    per = [np.sum(np.abs(full[weights == k]) ** 2) for k in range(w + 1)]
    # weight 1 is always filled before the first check (binomial.py:56-72)
    assert w >= 1 and all(sum(per[: k + 1]) <= max_l2 for k in range(1, w)) and max_l2 < sum(per)
    assert np.array_equal(G[weights <= w], full[weights <= w])
    assert abs(norm - sum(per)) < 1e-12


def test_oracle_global_cutoff(oracle_mod):
    from oracle import binomial as OB

    A, b, c = normalized_ket(3, 3)
    shape = (8, 8, 8)
    G, _ = OB.binomial(shape, A, b, c, None, 5)
    weights = np.sum(np.indices(shape), axis=0)
    full = oracle_mod.vanilla(shape, A, b, c)
    assert np.array_equal(G[weights <= 4], full[weights <= 4]) and not np.any(G[weights > 4])


# ----------------------------------------------------------------- GPU

CASES = [
    ((40,), None, 40, 1),
    ((12, 9), 0.9, 20, 2),
    ((10, 10), None, 19, 3),
    ((6, 5, 7), 0.99, 16, 4),
    ((5, 4, 6, 3), 0.999, 15, 5),
    ((8, 8, 8), None, 5, 6),
    ((30, 30, 30), 0.99999, 88, 8),
]


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
@pytest.mark.parametrize("shape,max_l2,gcut,seed", CASES, ids=lambda x: str(x))
def test_binomial_hip_vs_oracle(oracle_mod, shape, max_l2, gcut, seed):
    from oracle import binomial as OB

    from mrmustard_amd.lattice.strategies import binomial_hip

    A, b, c = normalized_ket(len(shape), seed)
    print(f"binomial shape={shape} max_l2={max_l2} gcut={gcut}", flush=True)
    G, norm = binomial_hip(shape, A, b, c, max_l2, gcut)
    ref, ref_norm = OB.binomial(shape, A, b, c, max_l2, gcut)
    assert np.array_equal(G.cpu().numpy(), ref)
    assert abs(norm - ref_norm) <= 1e-12 * max(1.0, ref_norm)


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
def test_binomial_manager_defaults(oracle_mod):
    """math.hermite_renormalized_binomial with None arguments uses
    AUTOSHAPE_PROBABILITY and sum(shape) - len(shape) + 1 (backend_numpy.py:414-421)."""
    from oracle import binomial as OB

    from mrmustard_amd import math, settings

    A, b, c = normalized_ket(2, 11)
    shape = (20, 20)
    G = math.hermite_renormalized_binomial(A, b, c, shape, None, None)
    ref, _ = OB.binomial(shape, A, b, c, settings.AUTOSHAPE_PROBABILITY, sum(shape) - len(shape) + 1)
    assert np.array_equal(G.cpu().numpy(), ref)


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
def test_binomial_errors():
    from mrmustard_amd.lattice.strategies import binomial_hip

    A, b, c = normalized_ket(2, 1)
    with pytest.raises(ValueError):
        binomial_hip((5, 5), A, b, c, 0.9, 0)
    with pytest.raises(ValueError):
        binomial_hip((5, 5, 5), A, b, c, 0.9, 4)
