"""Non-finite values: lattices that overflow, and cotangents holding inf/NaN.

The reference yields inf where a term overflows and NaN wherever such a value
meets a zero: its zero-weight terms A_ij * SQRT[0] * G[...] read wrapped-around
positions (core.py:101-103, 117-118), its VJP multiplies zero matrix entries
and SQRT[0] factors by G and by the cotangent (gradients.py:66-78), and
numba's complex / float division turns the other part of an infinite number
into NaN.  Parity here means the same values at every position, NaN and inf
included (inf signs too); finite VJP entries agree to 1e-12 (their sums run
in another order), non-finite ones exactly (a sum holding an inf or a NaN is
the same in every order).

CPU tests pin the oracle's non-finite behaviour to CPython's own complex
arithmetic on lists (tests/brute.py), which numba follows; GPU tests compare
every execution path of the library with the oracle's reference-order sums.
"""

from __future__ import annotations

import numpy as np
import pytest

from tests.brute import python_complex_stable, python_complex_vanilla, python_complex_vjp
from tests.conftest import gpu_available

TOL = 1e-12


def same_values(a, b) -> bool:
    """Equal values, NaN where the other has NaN (parts separately)."""
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.real, b.real, equal_nan=True) and np.array_equal(a.imag, b.imag, equal_nan=True)


def same_class_close(x, ref, tol=TOL) -> bool:
    """NaN and +-inf at the same parts, finite parts within tol (normwise)."""
    x, ref = np.asarray(x).ravel(), np.asarray(ref).ravel()
    xs = np.concatenate([x.real, x.imag])
    rs = np.concatenate([ref.real, ref.imag])
    if not np.array_equal(np.isnan(xs), np.isnan(rs)):
        return False
    inf = np.isinf(rs)
    if not np.array_equal(np.isinf(xs), inf) or not np.array_equal(xs[inf], rs[inf]):
        return False
    fin = np.isfinite(rs)
    if not fin.any():
        return True
    scale = max(np.max(np.abs(rs[fin])), 1e-300)
    return float(np.max(np.abs(xs[fin] - rs[fin]))) <= tol * scale


def blowup_triple(seed, D, sb, sA=1.0):
    """A triple whose lattice overflows part-way (inf and NaN mixed in)."""
    rng = np.random.default_rng(seed)
    A = (rng.normal(size=(D, D)) + 1j * rng.normal(size=(D, D))) * sA
    b = (rng.normal(size=D) + 1j * rng.normal(size=D)) * sb
    return A, b, complex(0.4, -0.3)


def tame_triple(seed, D):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(D, D)) + 1j * rng.normal(size=(D, D))
    A *= 0.45 / max(1.0, np.max(np.abs(np.linalg.eigvals(A))))
    b = 0.5 * (rng.normal(size=D) + 1j * rng.normal(size=D))
    return A, b, complex(rng.normal(), rng.normal())


def poisoned_cotangent(rng, shape, n_bad=3):
    g = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    bad = [complex(np.inf, 0.0), complex(np.inf, np.inf), complex(0.0, -np.inf), complex(np.nan, 0.0)]
    flat = g.reshape(-1)
    for k, pos in enumerate(rng.choice(flat.size, size=n_bad, replace=False)):
        flat[pos] = bad[k % len(bad)]
    return g


def has_mix(G) -> bool:
    return bool(np.isfinite(G).any() and np.isnan(G).any())


# ---------------------------------------------------------------- CPU: oracle


@pytest.mark.parametrize("shape,sb", [((9, 8), 1e60), ((5, 4, 4), 1e80), ((7,), 1e160), ((3, 3, 3, 3), 1e100), ((1, 6, 5), 1e70)],
                         ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple) else f"{v:g}")
def test_oracle_blowup_matches_python_complex(oracle_mod, shape, sb):
    A, b, c = blowup_triple(sum(shape), len(shape), sb, sA=1e40)
    G = oracle_mod.vanilla(shape, A, b, c)
    assert has_mix(G)
    assert same_values(G, python_complex_vanilla(A, b, c, shape))
    rng = np.random.default_rng(7)
    g = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    for x, r in zip(oracle_mod.vanilla_vjp(G, c, g), python_complex_vjp(G, c, g)):
        assert same_values(x, r)


@pytest.mark.parametrize("shape", [(7, 6, 5), (4, 9), (11,)], ids=lambda s: "x".join(map(str, s)))
def test_oracle_poisoned_cotangent_matches_python_complex(oracle_mod, shape):
    A, b, c = tame_triple(sum(shape), len(shape))
    G = oracle_mod.vanilla(shape, A, b, c)
    g = poisoned_cotangent(np.random.default_rng(sum(shape)), shape)
    for x, r in zip(oracle_mod.vanilla_vjp(G, c, g), python_complex_vjp(G, c, g)):
        assert same_values(x, r)


@pytest.mark.parametrize("shape,sb", [((9, 8), 1e60), ((5, 4, 4), 1e80), ((1, 6, 5), 1e70), ((3, 3, 3, 3), 1e100)],
                         ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple) else f"{v:g}")
def test_oracle_stable_blowup_matches_python_complex(oracle_mod, shape, sb):
    A, b, c = blowup_triple(sum(shape) + 1, len(shape), sb, sA=1e40)
    G = oracle_mod.stable(shape, A, b, c)
    assert has_mix(G)
    assert same_values(G, python_complex_stable(A, b, c, shape))


def test_zero_weight_terms_matter(oracle_mod):
    """The checks above see the zero-weight terms: without them (reading
    nothing where a digit is zero) the overflowing lattice comes out with
    fewer NaNs."""
    shape = (9, 8)
    A, b, c = blowup_triple(17, 2, 1e60, sA=1e40)
    ref = oracle_mod.vanilla(shape, A, b, c)
    Al = [[complex(x) for x in row] for row in A]
    bl = [complex(x) for x in b]
    G = np.zeros(shape, complex)
    G[0, 0] = c
    sq = np.sqrt(np.arange(20.0))
    for k0 in range(shape[0]):
        for k1 in range(shape[1]):
            if k0 == k1 == 0:
                continue
            if k0 > 0:
                v = bl[0] * complex(G[k0 - 1, k1])
                if k0 > 1:
                    v += Al[0][0] * sq[k0 - 1] * complex(G[k0 - 2, k1])
                if k1 > 0:
                    v += Al[0][1] * sq[k1] * complex(G[k0 - 1, k1 - 1])
                G[k0, k1] = v / sq[k0]
            else:
                v = bl[1] * complex(G[0, k1 - 1])
                if k1 > 1:
                    v += Al[1][1] * sq[k1 - 1] * complex(G[0, k1 - 2])
                G[0, k1] = v / sq[k1]
    assert not same_values(G, ref)
    assert np.isnan(ref).sum() > np.isnan(G).sum()


# ------------------------------------------------------------------ GPU paths

needs_gpu = pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")


def np_(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def st():
    import torch

    from mrmustard_amd.lattice import strategies

    torch.cuda.set_device(0)
    return strategies


# (shape, scale of b): LDS workgroup sweep, one-point-per-thread sweep,
# global slabs after an LDS face (a D = 2 and a D = 3 sub-triple's lattice,
# slabs two per launch; (6, 60, 60): a slab that fits in LDS, spread over all
# CUs for a single forward), and one-index chains
BLOWUP = [((60, 60), 1e6), ((30, 30, 30), 1e6), ((4, 80, 80), 1e4), ((3, 30, 30, 40), 1e4), ((6, 60, 60), 1e6), ((300,), 1e4), ((6, 5), 1e80)]


@pytest.mark.parametrize("shape,sb", BLOWUP, ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple) else f"{v:g}")
@pytest.mark.parametrize("mode", ["forward", "fused"])
@pytest.mark.gpu
@needs_gpu
def test_blowup_single(st, oracle_mod, shape, sb, mode):
    A, b, c = blowup_triple(sum(shape) + 3, len(shape), sb)
    ref = oracle_mod.vanilla(shape, A, b, c)
    assert has_mix(ref)
    print(f"blow-up {mode} shape={shape}", flush=True)
    if mode == "forward":
        assert same_values(np_(st.vanilla_hip(shape, A, b, c)), ref)
        return
    g = np.random.default_rng(1).normal(size=shape) + 1j * np.random.default_rng(2).normal(size=shape)
    G, dA, db, dc = st.vanilla_fused_hip(shape, A, b, c, g)
    assert same_values(np_(G), ref)
    for x, r in zip((dA, db, dc), oracle_mod.vanilla_vjp(ref, c, g)):
        assert same_class_close(np_(x), r)


def _batch(B, shape, bad, sb, seed):
    trip = [blowup_triple(seed + k, len(shape), sb) if k in bad else tame_triple(seed + k, len(shape)) for k in range(B)]
    A = np.stack([t[0] for t in trip])
    b = np.stack([t[1] for t in trip])
    c = np.array([t[2] for t in trip])
    return A, b, c


@pytest.mark.parametrize("shape,sb", [((30, 30, 30), 1e6), ((12,), 1e40), ((4, 80, 80), 1e4), ((20, 30), 1e8)],
                         ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple) else f"{v:g}")
@pytest.mark.gpu
@needs_gpu
def test_blowup_batch(st, oracle_mod, shape, sb):
    """Uniform batches with two overflowing triples among tame ones: forward,
    fused forward + VJP, and the stored-G VJP."""
    B = 5
    A, b, c = _batch(B, shape, {1, 3}, sb, sum(shape))
    ref = oracle_mod.vanilla_batch(shape, A, b, c)
    assert has_mix(ref[1])
    print(f"blow-up batch shape={shape}", flush=True)
    assert same_values(np_(st.vanilla_batch_hip(shape, A, b, c)), ref)
    g = np.random.default_rng(3).normal(size=ref.shape) + 1j * np.random.default_rng(4).normal(size=ref.shape)
    refs = [oracle_mod.vanilla_vjp(ref[k], c[k], g[k]) for k in range(B)]
    Gf, fA, fb, fc = st.vanilla_batch_fused_hip(shape, A, b, c, g)
    assert same_values(np_(Gf), ref)
    vA, vb, vc = st.vanilla_batch_vjp_hip(ref, c, g)
    for got in ((fA, fb, fc), (vA, vb, vc)):
        for k in range(B):
            for x, r in zip(got, refs[k]):
                assert same_class_close(np_(x)[k], r), (k, shape)


@pytest.mark.parametrize("shape", [(7, 6, 5), (30, 30, 30), (4, 80, 80), (2, 10, 10, 10, 10, 10), (11,)], ids=lambda s: "x".join(map(str, s)))
@pytest.mark.gpu
@needs_gpu
def test_poisoned_cotangent(st, oracle_mod, shape):
    """Finite G, a cotangent with inf and NaN entries: stored-G VJP (grouped
    and slab-by-slab kernels) and the fused sweep."""
    A, b, c = tame_triple(sum(shape) + 9, len(shape))
    G = oracle_mod.vanilla(shape, A, b, c)
    g = poisoned_cotangent(np.random.default_rng(sum(shape)), shape)
    refs = oracle_mod.vanilla_vjp(G, c, g)
    print(f"poisoned cotangent shape={shape}", flush=True)
    for x, r in zip(st.vanilla_vjp_hip(G, c, g), refs):
        assert same_class_close(np_(x), r)
    Gf, dA, db, dc = st.vanilla_fused_hip(shape, A, b, c, g)
    assert np.array_equal(np_(Gf), G)
    for x, r in zip((dA, db, dc), refs):
        assert same_class_close(np_(x), r)


@pytest.mark.parametrize("case", ["blowup-4x80x80", "blowup-3x30x30x40", "poisoned-4x80x80"])
@pytest.mark.gpu
@needs_gpu
def test_recompute_nonfinite(st, oracle_mod, case):
    """The windowed recompute VJP (slabs too large for one workgroup's LDS) on
    an overflowing lattice and with a cotangent holding inf and NaN: its
    non-finite dL/dc sends the call through the stored-lattice path, whose
    values are the reference's."""
    kind, dims = case.split("-")
    shape = tuple(int(x) for x in dims.split("x"))
    rng = np.random.default_rng(sum(shape) + 11)
    if kind == "blowup":
        A, b, c = blowup_triple(sum(shape) + 3, len(shape), 1e4)
        g = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    else:
        A, b, c = tame_triple(sum(shape) + 9, len(shape))
        g = poisoned_cotangent(rng, shape)
    G = oracle_mod.vanilla(shape, A, b, c)
    assert has_mix(G) if kind == "blowup" else np.isfinite(G).all()
    refs = oracle_mod.vanilla_vjp(G, c, g)
    print(f"recompute non-finite {case}", flush=True)
    for x, r in zip(st.vanilla_vjp_recompute_hip(shape, A, b, c, g), refs):
        assert same_class_close(np_(x), r)


@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.gpu
@needs_gpu
def test_blowup_plan(oracle_mod, st, fused):
    """The work-queue plan (workgroup, wave and lane teams) with overflowing
    triples of every team size."""
    from mrmustard_amd.batch import HermitePlan

    shapes = [(30, 30, 30), (60, 60), (6, 5), (8, 7), (300,), (12,), (20, 30), (5, 4, 3)]
    scales = [1e6, 1e6, 1e80, 1e60, 1e4, 1e40, 1e8, 1e100]
    triples = []
    for k, (s, sb) in enumerate(zip(shapes, scales)):
        triples.append(blowup_triple(100 + k, len(s), sb) if k % 2 == 0 else tame_triple(100 + k, len(s)))
    rng = np.random.default_rng(8)
    gs = [rng.normal(size=s) + 1j * rng.normal(size=s) for s in shapes] if fused else None
    plan = HermitePlan(shapes, fused=fused)
    A, b, c = plan.pack(triples)
    out = plan.run(A, b, c, dLdG=plan.pack_amplitudes(gs) if fused else None)
    Gs_ref, grads_ref = oracle_mod.uneven(triples, shapes, gs)
    G = out[0] if fused else out
    print(f"blow-up plan fused={fused}", flush=True)
    for k, (x, r) in enumerate(zip(plan.split(G), Gs_ref)):
        assert same_values(np_(x), r), (k, shapes[k])
    if fused:
        for k, (got, ref) in enumerate(zip(plan.split_grads(*out[1:]), grads_ref)):
            for x, r in zip(got, ref):
                assert same_class_close(np_(x), r), (k, shapes[k])


@pytest.mark.parametrize("shape,sb", [((60, 60), 1e6), ((30, 30, 30), 1e6), ((20, 20), 1e10), ((600, 500), 1e3), ((40, 40, 40), 1e4),
                                      ((40, 40, 40), 1e8), ((4, 4, 4, 4, 40), 1e8), ((3500, 9), 1e2)],
                         ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple) else f"{v:g}")
@pytest.mark.gpu
@needs_gpu
def test_stable_blowup(st, oracle_mod, shape, sb):
    """The all-pivot average on overflowing lattices: one workgroup per triple
    (few rows of the last axis), and the per-wavefront sweep ((40, 40, 40):
    one launch per wavefront, overflowing mid-sweep at 1e4 and in the first
    wavefronts at 1e8; (4, 4, 4, 4, 40) at D = 5; (3500, 9): the first and
    last wavefronts in one workgroup each, overflowing in the first), then the
    reference's sequential sweep where a non-finite value appeared."""
    A, b, c = blowup_triple(sum(shape) + 5, len(shape), sb)
    ref = oracle_mod.stable(shape, A, b, c)
    assert has_mix(ref)
    print(f"stable blow-up shape={shape}", flush=True)
    assert same_values(np_(st.stable_hip(shape, A, b, c)), ref)


@pytest.mark.gpu
@needs_gpu
def test_stable_blowup_batch(st, oracle_mod):
    shape = (30, 30, 30)
    A, b, c = _batch(4, shape, {2}, 1e6, 77)
    ref = oracle_mod.vanilla_batch(shape, A, b, c, stable=True)
    assert has_mix(ref[2])
    print("stable blow-up batch", flush=True)
    assert same_values(np_(st.vanilla_batch_hip(shape, A, b, c, stable=True)), ref)
