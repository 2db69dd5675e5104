"""Autotuner for work-queue plans, keyed by triple class.

A plan's launch configuration is the threads per workgroup team and the
largest axis-0 slab handed to a single wave instead of a workgroup.  The best
choice depends on the mix of triples, so the tuner times each *class* of
triple once -- the number of indices D, the size class ceil(log2) of the slab
prod(shape[1:]), forward or fused -- on a synthetic batch of that class under
every team option (one wave per triple; one workgroup of 256, 512 or 768
threads), and keeps the cost per triple in a table.  A batch's configuration
is then read off the table: the threads and the wave threshold that minimise
the sum of the per-triple costs of its classes.  Only classes missing from
the table are timed, so a new batch of known classes never tunes on the call
path.

The table ships with the package (``autotune_table.json``, measured on an
MI355X with ``python -m mrmustard_amd.autotune``) and is extended by a user
cache (``$MM_AUTOTUNE_CACHE`` or ``~/.cache/mrmustard_amd/autotune_classes.json``).
One-index chains (D = 1) always run one per lane and need no entry.
"""

from __future__ import annotations

import json
import math
import os

import numpy as np
import torch

from .batch import HermitePlan

_TABLE: dict[str, dict] | None = None
_PACKAGE_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_table.json")
SYNTH_TRIPLES = 512
SYNTH_MAX_AMPS = 20_000_000
AXIS0 = 32  # axis-0 length of the synthetic triples of a class


def _cache_path() -> str:
    return os.environ.get(
        "MM_AUTOTUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "mrmustard_amd", "autotune_classes.json")
    )


def _arch() -> str:
    dev = torch.cuda.get_device_properties(torch.cuda.current_device())
    return getattr(dev, "gcnArchName", "gpu").split(":")[0]


def slab_class(shape) -> int:
    """ceil(log2) of the axis-0 slab prod(shape[1:])."""
    return int(math.ceil(math.log2(max(1, int(np.prod(shape[1:]))))))


def class_key(D: int, sc: int, fused: bool, arch: str | None = None) -> str:
    return f"{arch or _arch()}/{'fused' if fused else 'fwd'}/D{D}/s{sc}"


def team_threads(fused: bool, max_d: int) -> list[int]:
    """Workgroup sizes a plan accepts (the LDS kernel's launch bound)."""
    if fused:
        return [256] if max_d > 3 else [256, 512]
    return [256, 512] if max_d > 3 else [256, 512, 768]


def _load_table() -> dict:
    global _TABLE
    if _TABLE is None:
        table = {}
        for path in (_PACKAGE_TABLE, _cache_path()):
            if os.path.exists(path):
                try:
                    table.update(json.load(open(path)))
                except (OSError, ValueError):
                    pass
        _TABLE = table
    return _TABLE


def _save_user(entries: dict) -> None:
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        merged = {}
        if os.path.exists(path):
            try:
                merged = json.load(open(path))
            except (OSError, ValueError):
                merged = {}
        merged.update(entries)
        json.dump(merged, open(path, "w"), indent=0, sort_keys=True)
    except OSError:
        pass


def representative_shape(D: int, sc: int) -> tuple:
    """A shape of D indices whose slab has about 2^sc points."""
    target = 2.0**sc
    inner = [max(1, int(round(target ** (1.0 / (D - 1)))))] * (D - 1)
    # adjust the last inner axis towards the target product
    rest = int(np.prod(inner[:-1])) if D > 2 else 1
    inner[-1] = max(1, int(round(target / rest)))
    return (AXIS0, *inner)


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
    # well-conditioned triples of the right sizes (values do not change the timing)
    A = np.zeros(plan.a_elems, np.complex128)
    for k, D in enumerate(plan.dims.tolist()):
        a0 = int(plan.a_off[k])
        A[a0 : a0 + D * D] = (0.3 * np.eye(D)).reshape(-1)
    b = 0.3 * (rng.normal(size=plan.b_elems) + 1j * rng.normal(size=plan.b_elems))
    c = np.ones(len(plan.shapes), np.complex128)
    args = [torch.as_tensor(A, device=dev), torch.as_tensor(b, device=dev), torch.as_tensor(c, device=dev)]
    if fused:
        args.append(torch.ones(plan.n_amplitudes, dtype=torch.complex128, device=dev))
    return args


def tune_class(D: int, sc: int, fused: bool) -> dict:
    """Time one class under every team option: microseconds per triple."""
    shape = representative_shape(D, sc)
    n = int(max(16, min(SYNTH_TRIPLES, SYNTH_MAX_AMPS // int(np.prod(shape)))))
    shapes = [shape] * n
    slab = int(np.prod(shape[1:]))
    entry = {"shape": list(shape), "triples": n, "wave": None, "block": {}}
    options = [("wave", {"block_threads": 256, "wave_max_slab": slab})]
    options += [(str(t), {"block_threads": t, "wave_max_slab": -1}) for t in team_threads(fused, D)]
    for name, tune in options:
        try:
            plan = HermitePlan(shapes, fused=fused, tune=tune)
        except (ValueError, RuntimeError):
            continue
        d = plan.describe()
        if name == "wave" and d["wave_triples"] != n:
            plan.close()
            continue
        us = 1e3 * _time_plan(plan, _synthetic_inputs(plan, fused)) / n
        plan.close()
        if name == "wave":
            entry["wave"] = us
        else:
            entry["block"][name] = us
    return entry


def choose(shapes, fused: bool, table: dict, arch: str | None = None) -> dict:
    """The (threads, wave threshold) minimising the summed class costs."""
    classes: dict[tuple, int] = {}
    max_d = 1
    for s in shapes:
        if len(s) == 1:
            continue
        max_d = max(max_d, len(s))
        key = (len(s), slab_class(s))
        classes[key] = classes.get(key, 0) + 1
    if not classes:
        return {"block_threads": 256, "wave_max_slab": -1, "blocks_per_cu": 0}
    thresholds = sorted({-1} | {2**sc for (_, sc) in classes})
    best, best_cost = None, math.inf
    for t in team_threads(fused, max_d):
        for thr in thresholds:
            cost = 0.0
            for (D, sc), cnt in classes.items():
                e = table[class_key(D, sc, fused, arch)]
                use_wave = thr > 0 and 2**sc <= thr and e["wave"] is not None
                c = e["wave"] if use_wave else e["block"].get(str(t))
                if c is None:  # not measured for this team size: the class's best, penalised
                    c = min(e["block"].values()) * 1.5 if e["block"] else math.inf
                cost += cnt * c
            if cost < best_cost:
                best, best_cost = {"block_threads": t, "wave_max_slab": thr, "blocks_per_cu": 0}, cost
    return best


def tuned_plan(shapes, fused: bool = False, retune: bool = False):
    """Return (plan, info) with the launch configuration the class table
    gives for ``shapes``; classes missing from the table are timed first."""
    shapes = [tuple(int(x) for x in s) for s in shapes]
    table = _load_table()
    needed = sorted({(len(s), slab_class(s)) for s in shapes if len(s) > 1})
    new = {}
    for D, sc in needed:
        key = class_key(D, sc, fused)
        if retune or key not in table:
            new[key] = tune_class(D, sc, fused)
    if new:
        table.update(new)
        _save_user(new)
    tune = choose(shapes, fused, table)
    plan = HermitePlan(shapes, fused=fused, tune=tune)
    info = {"tune": tune, "cached": not new, "classes": len(needed), "classes_timed": len(new), "describe": plan.describe()}
    return plan, info


def write_package_table(path: str = _PACKAGE_TABLE, D_max: int = 3, sc_max: int = 12) -> str:
    """Time every class with 2 <= D <= D_max and slab classes up to sc_max,
    forward and fused, and write the table (by default the package's)."""
    table = {}
    for fused in (False, True):
        for D in range(2, D_max + 1):
            for sc in range(1, sc_max + 1):
                if D == 2 and 2**sc > 4096:
                    continue
                table[class_key(D, sc, fused)] = tune_class(D, sc, fused)
                print(class_key(D, sc, fused), table[class_key(D, sc, fused)], flush=True)
    json.dump(table, open(path, "w"), indent=0, sort_keys=True)
    return path


if __name__ == "__main__":
    import sys

    torch.cuda.set_device(0)
    print(write_package_table(*sys.argv[1:2]))
