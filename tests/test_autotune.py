"""The class-keyed tuning table (CPU: no timing, the shipped table only)."""

from __future__ import annotations

import json

import numpy as np
import pytest

torch = pytest.importorskip("torch")


def test_shipped_table_covers_the_uneven_workloads():
    from mrmustard_amd import autotune as T
    from mrmustard_amd import workloads as W

    table = json.load(open(T._PACKAGE_TABLE))
    for seed in (2, 102, 7):
        _, shapes = W.uneven_kets(seed=seed)
        for fused in (False, True):
            for s in shapes:
                if len(s) > 1:
                    assert T.class_key(len(s), T.slab_class(s), fused, "gfx950") in table, s


def test_choose_minimises_the_class_costs():
    from mrmustard_amd import autotune as T

    # a made-up table: waves win for slabs up to 64 points, 512-thread teams above
    table = {}
    for sc in range(1, 13):
        wave = 1.0 if sc <= 6 else 50.0
        table[T.class_key(3, sc, False, "x")] = {"wave": wave, "block": {"256": 3.0, "512": 2.0 + (sc <= 6), "1024": 4.0}}
    shapes = [(10, 4, 4)] * 50 + [(10, 30, 30)] * 20
    tune = T.choose(shapes, False, table, "x")
    assert tune["block_threads"] == 512 and tune["wave_max_slab"] == 16
    # all classes above the wave threshold: no waves at all
    assert T.choose([(10, 30, 30)] * 5, False, table, "x")["wave_max_slab"] == -1


def test_representative_shapes_hit_their_class():
    from mrmustard_amd import autotune as T

    for D in (2, 3, 4):
        for sc in range(1, 13):
            s = T.representative_shape(D, sc)
            assert len(s) == D
            assert abs(T.slab_class(s) - sc) <= 1, (D, sc, s)
    assert np.all([T.slab_class(T.representative_shape(3, sc)) == sc for sc in (4, 6, 8, 10, 12)])
