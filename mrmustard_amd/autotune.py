"""Autotuner for work-queue plans.

For a given list of shapes it builds the plan under a small set of launch
configurations (threads per workgroup, the largest axis-0 slab handed to a
single wave instead of a workgroup, workgroups per compute unit), times each
on the GPU with HIP events, and keeps the fastest.  Results are cached per
workload signature in memory and in a JSON file
(``$MM_AUTOTUNE_CACHE`` or ``~/.cache/mrmustard_amd/autotune.json``).
"""

from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import torch

from .batch import HermitePlan

_MEM: dict[str, dict] = {}


def _cache_path() -> str:
    return os.environ.get("MM_AUTOTUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "mrmustard_amd", "autotune.json"))


def signature(shapes, fused: bool) -> str:
    h = hashlib.sha1()
    h.update(b"fused" if fused else b"fwd")
    for s in shapes:
        h.update(np.asarray(s, dtype=np.int64).tobytes())
        h.update(b"|")
    dev = torch.cuda.get_device_properties(torch.cuda.current_device())
    h.update(getattr(dev, "gcnArchName", "gpu").encode())
    return h.hexdigest()[:20]


def candidates(fused: bool, max_d: int):
    threads = [256, 512] if fused else [256, 512, 1024]
    if fused and max_d > 3:
        threads = [256]
    out = []
    for t in threads:
        for wave in (-1, 64, 128):
            out.append({"block_threads": t, "wave_max_slab": wave, "blocks_per_cu": 0})
    return out


def _time_plan(plan: HermitePlan, args, reps: int = 3) -> float:
    stream = torch.cuda.current_stream()
    plan.run(*args)  # warm-up
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        plan.run(*args)
        e1.record(stream)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    return float(np.median(times))


def _synthetic_inputs(plan: HermitePlan, fused: bool):
    dev = torch.device("cuda", torch.cuda.current_device())
    rng = np.random.default_rng(0)
    # well-conditioned random triples of the right sizes (values do not affect timing)
    A = np.zeros(plan.a_elems, np.complex128)
    for k, D in enumerate(plan.dims.tolist()):
        a0 = int(plan.a_off[k])
        A[a0 : a0 + D * D] = (0.3 * np.eye(D)).reshape(-1)
    b = 0.3 * (rng.normal(size=plan.b_elems) + 1j * rng.normal(size=plan.b_elems))
    c = np.ones(len(plan.shapes), np.complex128)
    args = [torch.as_tensor(A, device=dev), torch.as_tensor(b, device=dev), torch.as_tensor(c, device=dev)]
    if fused:
        g = torch.ones(plan.n_amplitudes, dtype=torch.complex128, device=dev)
        args.append(g)
    return args


def tuned_plan(shapes, fused: bool = False, retune: bool = False):
    """Return (plan, info) with the fastest launch configuration for `shapes`."""
    shapes = [tuple(int(x) for x in s) for s in shapes]
    sig = signature(shapes, fused)
    path = _cache_path()
    if not retune and sig not in _MEM and os.path.exists(path):
        try:
            _MEM.update(json.load(open(path)))
        except (OSError, ValueError):
            pass
    if not retune and sig in _MEM:
        best = _MEM[sig]
        return HermitePlan(shapes, fused=fused, tune=best["tune"]), dict(best, cached=True)
    max_d = max(len(s) for s in shapes) if shapes else 1
    results = []
    best_plan, best = None, None
    for tune in candidates(fused, max_d):
        try:
            plan = HermitePlan(shapes, fused=fused, tune=tune)
        except ValueError:
            continue
        args = _synthetic_inputs(plan, fused)
        ms = _time_plan(plan, args)
        results.append({"tune": tune, "ms": ms})
        if best is None or ms < best["ms"]:
            if best_plan is not None:
                best_plan.close()
            best_plan, best = plan, {"tune": tune, "ms": ms}
        else:
            plan.close()
    info = dict(best, tried=len(results), cached=False, describe=best_plan.describe())
    _MEM[sig] = {"tune": best["tune"], "ms": best["ms"]}
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        merged = {}
        if os.path.exists(path):
            try:
                merged = json.load(open(path))
            except (OSError, ValueError):
                merged = {}
        merged[sig] = _MEM[sig]
        json.dump(merged, open(path, "w"))
    except OSError:
        pass
    return best_plan, info
