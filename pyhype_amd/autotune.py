"""Launch-shape autotuner for the stage kernels.

Fused variant: the work item is a 60-column strip of one block over ``rows_per_item`` rows; each
item recomputes one reconstruction row above and below its range, so short items waste work while
long items leave SIMDs idle on small meshes.  Shared-halo variant: workgroups of 4 waves with
``rows_per_item`` rows each exchange their boundary rows through LDS.  Split variant: a reconstruction kernel with no row
halo (``rows_per_item`` rows per item) and a y-face kernel (``rows_b`` rows per item).
Element-wise variant: one lane per cell and variable for the reconstruction, one per face for
the fluxes and one per cell for the update (no shape parameters).  Paired variant: the fused
kernel's items with two waves each, one per pair of variables (``rows_per_item`` rows).  Variant
5: the split stage with its y-face and update kernels fused (``rows_b`` cell rows per lane);
variant 6: the same with one face per wave (``rows_b`` rows and ``rows_b`` + 1 waves per workgroup);
variant 7: variant 6 with the reconstruction over row-major positions (62 per wave, no shape
parameter of its own); variant 8: that reconstruction with the y-faces and update over flat cell
ranges (``rows_b`` + 1 waves per workgroup, at least ``rows_per_item`` faces per lane).
``autotune`` times whole steps (hipGraph replay) for a set of candidate shapes and keeps the
fastest; the fastest split or element-wise shape is then also timed with its work lists dealt to the
XCDs in chunks of ``XCD_CHUNKS`` items (``xcd_chunk``; 0 = one contiguous range per XCD).  Results are
cached in a JSON file keyed by a signature of the workload (geometry kind, flux, stages, block
shapes, device).  The state, time and step counter of the plan are restored before every trial
and after tuning.
"""
from __future__ import annotations

import json
import os
import time
from collections import Counter
from pathlib import Path
from typing import Dict, Optional, Sequence

CACHE = Path(os.environ.get("PYHYPE_AMD_TUNING", Path(__file__).resolve().parent / "tuning.json"))
ROWS = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 64)


def signature(plan) -> str:
    import torch

    kinds = "cart" if all(e.cartesian for e in plan.blocks) else (
        "gen" if not any(e.cartesian for e in plan.blocks) else "mixed")
    shapes = Counter((e.nx, e.ny) for e in plan.blocks)
    shp = ",".join(f"{n}x{a}x{b}" for (a, b), n in sorted(shapes.items()))
    dev = getattr(torch.cuda.get_device_properties(0), "gcnArchName", "gpu").split(":")[0] \
        if torch.cuda.is_available() else "none"
    lim = getattr(plan, "limiter", "Venkatakrishnan")
    flux = plan.flux if lim == "Venkatakrishnan" else f"{plan.flux}/{lim}"
    if getattr(plan, "primitive", False):
        flux += "/prim"
    if getattr(plan, "quadrature_points", 1) > 1:
        flux += f"/q{plan.quadrature_points}"
    return f"{kinds}|{flux}|{len(plan.table)}|{shp}|{dev}"


def load_cache(path: Path = CACHE) -> Dict[str, dict]:
    try:
        return json.loads(Path(path).read_text())
    except (OSError, ValueError):
        return {}


def lookup(plan, path: Path = CACHE) -> Optional[dict]:
    return load_cache(path).get(signature(plan))


MIN_TRIAL_S = 0.01  # a timed trial lasts at least this long (small meshes: more steps)


def _time_steps(plan, steps: int) -> float:
    """Seconds per step over `steps` graph-replayed steps, or over as many more as make the trial
    last MIN_TRIAL_S (on small meshes a few steps take well under a millisecond, where launch and
    synchronisation jitter would decide the ranking)."""
    import torch

    plan.run(1, use_graph=True, check=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.run(steps, use_graph=True, check=False)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    if el < MIN_TRIAL_S:
        steps = int(steps * min(64.0, MIN_TRIAL_S / max(el, 1e-6)) + 1)
        t0 = time.perf_counter()
        plan.run(steps, use_graph=True, check=False)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
    return el / steps


def candidates(rows: Sequence[int] = ROWS, grids: Sequence[int] = (0, 1024, 2048, 3072), waves: Sequence[int] = (2, 3),
               split_rows: Sequence[int] = (1, 2, 3, 4, 8), split_rows_b: Sequence[int] = (1, 2, 4, 8, 16, 32),
               split_waves: Sequence[int] = (3, 4), shared_rows: Sequence[int] = (1, 2, 3, 4, 6),
               paired_rows: Sequence[int] = (1, 2, 3, 4, 6, 8, 12)):
    """The default search space: fused shapes (one workgroup per work item, grid 0, or a persistent
    grid of `grids` workgroups that stride over the work list), shared-halo shapes (rows per wave),
    split shapes, the element-wise stage, paired shapes (rows per item), and persistent grids of the
    split stages' kernels."""
    out = [{"variant": 0, "rows_per_item": r, "grid_waves": g, "waves_per_simd": w, "rows_b": 0}
           for w in waves for g in grids for r in rows]
    out += [{"variant": 2, "rows_per_item": r, "grid_waves": 0, "waves_per_simd": 2, "rows_b": 0}
            for r in shared_rows]
    out += [{"variant": 1, "rows_per_item": r, "grid_waves": 0, "waves_per_simd": w, "rows_b": rb}
            for w in split_waves for r in split_rows for rb in split_rows_b]
    out += [{"variant": 3, "rows_per_item": 8, "grid_waves": 0, "waves_per_simd": 2, "rows_b": 0}]
    out += [{"variant": 4, "rows_per_item": r, "grid_waves": 0, "waves_per_simd": w, "rows_b": 0}
            for w in (2, 3) for r in paired_rows]
    out += [{"variant": 5, "rows_per_item": r, "grid_waves": 0, "waves_per_simd": w, "rows_b": rb}
            for w in split_waves for r in (1, 2) for rb in (1, 2, 3, 4, 8)]
    out += [{"variant": 6, "rows_per_item": r, "grid_waves": 0, "waves_per_simd": w, "rows_b": rb}
            for w in split_waves for r in (1, 2) for rb in (1, 2, 3, 4, 7)]
    out += [{"variant": 7, "rows_per_item": 1, "grid_waves": 0, "waves_per_simd": w, "rows_b": rb}
            for w in split_waves for rb in (1, 2, 3, 4, 7)]
    out += [{"variant": 8, "rows_per_item": r, "grid_waves": 0, "waves_per_simd": w, "rows_b": rb}
            for w in split_waves for r in (1, 2) for rb in (7, 15)]
    out += [{"variant": v, "rows_per_item": r, "grid_waves": g, "waves_per_simd": 3, "rows_b": rb}
            for v, r, rb in ((6, 1, 3), (7, 1, 3), (8, 1, 7)) for g in grids if g]
    return out


XCD_CHUNKS = (1, 4, 16, 64)   # work-list orders tried for the best split or element-wise shape
SHAPE_KEYS = ("variant", "rows_per_item", "rows_b", "grid_waves", "waves_per_simd", "xcd_chunk")


def _apply(plan, c):
    plan.set_launch(c["rows_per_item"], c["grid_waves"], c["waves_per_simd"], variant=c.get("variant", 0),
                    rows_b=c.get("rows_b", 0), xcd_chunk=c.get("xcd_chunk", 0))


def autotune(plan, steps: int = 10, shapes: Optional[Sequence[dict]] = None, path: Optional[Path] = CACHE,
             verbose: bool = False) -> dict:
    """Time candidate launch shapes (``candidates()`` by default) on the plan's current state;
    apply and cache the best.  Returns the best shape, its ms/step and all trials."""
    import numpy as np

    if shapes is None:
        shapes = candidates()
        if getattr(plan, "quadrature_points", 1) > 1:   # Gauss faces run the split stages only
            shapes = [c for c in shapes if c["variant"] in (1, 5, 6)]
    U_saved = plan.U.clone()
    _, step0, t0, _ = plan.status()
    trials, skipped = [], []
    best = None

    def trial(c):
        try:
            _apply(plan, c)
        except ValueError as e:   # a shape the library refuses for this mesh (e.g. LDS ranges)
            skipped.append(dict(c, reason=str(e)))
            return
        plan.U.copy_(U_saved)
        plan.reset(t0, step0)
        ms = _time_steps(plan, steps) * 1e3
        trials.append(dict(c, ms_per_step=ms))
        if verbose:
            print(f"autotune {c}: {ms:.4f} ms/step", flush=True)

    try:
        for c in shapes:
            trial(c)
        if not trials:
            raise ValueError("autotune: the library refused every candidate shape")
        split = [t for t in trials if t["variant"] != 0 and t["grid_waves"] == 0 and not t.get("xcd_chunk", 0)]
        if split:   # the fastest split or element-wise shape with its work lists dealt to the XCDs
            b = min(split, key=lambda t: t["ms_per_step"])
            for k in XCD_CHUNKS:
                trial(dict({key: b[key] for key in SHAPE_KEYS[:5]}, xcd_chunk=k))
        best = min(trials, key=lambda t: t["ms_per_step"])
        _apply(plan, best)
    finally:
        # the timed trials advanced the state: always hand the plan back as it was
        plan.U.copy_(U_saved)
        plan.reset(np.asarray(t0), step0)
    result = dict(best, trials=trials, skipped=skipped)
    if path is not None:
        cache = load_cache(path)
        entry = {k: best.get(k, 0) for k in SHAPE_KEYS}
        entry["ms_per_step"] = best["ms_per_step"]
        cache[signature(plan)] = entry
        Path(path).write_text(json.dumps(cache, indent=1, sort_keys=True))
    return result


def apply_cached(plan, path: Path = CACHE) -> bool:
    """Apply the cached shape for this workload if there is one."""
    hit = lookup(plan, path)
    if hit is None:
        return False
    _apply(plan, {"rows_per_item": int(hit["rows_per_item"]), "grid_waves": int(hit["grid_waves"]),
                  "waves_per_simd": int(hit.get("waves_per_simd", 2)), "variant": int(hit.get("variant", 0)),
                  "rows_b": int(hit.get("rows_b", 0)), "xcd_chunk": int(hit.get("xcd_chunk", 0))})
    return True
