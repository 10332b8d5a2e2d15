"""Host logic of the launch-shape autotuner (no GPU): candidate shapes within the limits the C
side accepts, workload signatures, the cache round trip, the Gauss-face filter, and the committed
tuning cache covering the tuned workloads."""
import json
from types import SimpleNamespace as NS

import pytest
import torch

from pyhype_amd import autotune as AT

YU_MAX_ROWS = 7          # one-y-face-per-wave stage: rows per workgroup
FLAT_MAX_WAVES = 16      # flat y-face ranges: waves per workgroup


def fake_plan(kinds, shapes, flux="Roe", stages=4, limiter="Venkatakrishnan", prim=False, nq=1):
    blocks = [NS(cartesian=c, nx=nx, ny=ny) for c, (nx, ny) in zip(kinds, shapes)]
    return NS(blocks=blocks, flux=flux, table=[None] * stages, limiter=limiter, primitive=prim,
              quadrature_points=nq)


def _strip_device(sig):
    return sig.rsplit("|", 1)[0]


def test_candidates_cover_every_variant_within_the_native_limits():
    cands = AT.candidates()
    assert {c["variant"] for c in cands} == set(range(9))
    for c in cands:
        assert set(c) == {"variant", "rows_per_item", "grid_waves", "waves_per_simd", "rows_b"}
        assert c["rows_per_item"] >= 1 and c["waves_per_simd"] in (2, 3, 4) and c["grid_waves"] >= 0
        if c["variant"] in (6, 7):
            assert 1 <= c["rows_b"] <= YU_MAX_ROWS
        if c["variant"] == 8:
            assert 1 <= c["rows_b"] <= FLAT_MAX_WAVES - 1
    assert len({json.dumps(c, sort_keys=True) for c in cands}) == len(cands)   # no duplicate trials


@pytest.mark.parametrize("plan,key", [
    (fake_plan([False] * 8, [(150, 150)] * 8), "gen|Roe|4|8x150x150"),
    (fake_plan([True] * 5, [(500, 500)] * 5, flux="HLLL", stages=2), "cart|HLLL|2|5x500x500"),
    (fake_plan([True, False, False, False], [(50, 50)] * 4, flux="HLLL", stages=2, prim=True),
     "mixed|HLLL/prim|2|4x50x50"),
    (fake_plan([True, True], [(64, 64), (32, 48)], limiter="VanLeer", nq=3), "cart|Roe/VanLeer/q3|4|1x32x48,1x64x64"),
])
def test_signature_names_geometry_scheme_stages_and_shapes(plan, key):
    assert _strip_device(AT.signature(plan)) == key


def test_apply_cached_sets_the_cached_launch(tmp_path):
    plan = fake_plan([False] * 2, [(40, 40)] * 2)
    calls = []
    plan.set_launch = lambda *a, **k: calls.append((a, k))
    path = tmp_path / "tuning.json"
    assert AT.apply_cached(plan, path) is False and not calls
    path.write_text(json.dumps({AT.signature(plan): {"variant": 8, "rows_per_item": 2, "rows_b": 7,
                                                     "grid_waves": 0, "waves_per_simd": 3, "ms_per_step": 1.0}}))
    assert AT.apply_cached(plan, path) is True
    assert calls == [((2, 0, 3), {"variant": 8, "rows_b": 7, "xcd_chunk": 0})]
    calls.clear()
    path.write_text(json.dumps({AT.signature(plan): {"variant": 7, "rows_per_item": 1, "rows_b": 3, "grid_waves": 0,
                                                     "waves_per_simd": 3, "xcd_chunk": 4, "ms_per_step": 1.0}}))
    assert AT.apply_cached(plan, path) is True
    assert calls == [((1, 0, 3), {"variant": 7, "rows_b": 3, "xcd_chunk": 4})]


def test_autotune_keeps_the_fastest_and_gauss_faces_try_split_stages_only(tmp_path, monkeypatch):
    plan = fake_plan([True], [(30, 30)], nq=2)
    plan.U = torch.zeros(8, dtype=torch.float64)
    state = {"launch": None, "resets": 0}
    plan.status = lambda: (None, 5, 0.25, None)

    def reset(t0, step0):
        assert step0 == 5 and float(t0) == 0.25
        state["resets"] += 1

    plan.reset = reset
    plan.set_launch = lambda r, g, w, variant=0, rows_b=0, xcd_chunk=0: state.update(launch=(variant, r, w, rows_b))
    # a made-up timing: variant 5 with rows_b 2 is fastest
    monkeypatch.setattr(AT, "_time_steps", lambda p, steps: 1.0 - 0.5 * (state["launch"][0] == 5)
                        - 0.1 * (state["launch"][3] == 2))
    path = tmp_path / "tuning.json"
    best = AT.autotune(plan, steps=1, path=path)
    assert {t["variant"] for t in best["trials"]} == {1, 5, 6}
    assert best["variant"] == 5 and best["rows_b"] == 2
    assert state["launch"][0] == 5 and state["launch"][3] == 2   # the winner is applied
    cached = json.loads(path.read_text())[AT.signature(plan)]
    assert cached["variant"] == 5 and cached["rows_b"] == 2


def test_autotune_tries_xcd_orders_for_the_best_split_shape(tmp_path, monkeypatch):
    """After the shape trials, the fastest split shape is timed again with its work lists dealt to
    the XCDs in chunks of 1, 4 and 16 items, also when a fused shape won the shape trials; the
    fastest of all trials is applied and cached with its order."""
    plan = fake_plan([True] * 5, [(500, 500)] * 5, flux="HLLL", stages=2)
    plan.U = torch.zeros(8, dtype=torch.float64)
    state = {"launch": None}
    plan.status = lambda: (None, 0, 0.0, None)
    plan.reset = lambda t0, step0: None
    plan.set_launch = lambda r, g, w, variant=0, rows_b=0, xcd_chunk=0: state.update(
        launch=(variant, r, g, w, rows_b, xcd_chunk))
    # made-up timings: the fused variant wins the shape trials (0.55), variant 7 with rows_b 3 is the
    # best split shape (0.6) and wins overall when dealt in chunks of 4 items (0.4)
    monkeypatch.setattr(AT, "_time_steps", lambda p, steps: 1.0 - 0.45 * (state["launch"][0] == 0)
                        - 0.3 * (state["launch"][0] == 7) - 0.1 * (state["launch"][4] == 3)
                        - 0.2 * (state["launch"][5] == 4))
    path = tmp_path / "tuning.json"
    best = AT.autotune(plan, steps=1, path=path)
    orders = [t for t in best["trials"] if t.get("xcd_chunk")]
    assert sorted(t["xcd_chunk"] for t in orders) == list(AT.XCD_CHUNKS)
    assert all(t["variant"] == 7 and t["rows_b"] == 3 for t in orders)
    assert best["variant"] == 7 and best["xcd_chunk"] == 4
    assert state["launch"][0] == 7 and state["launch"][5] == 4          # the winner is applied
    assert json.loads(path.read_text())[AT.signature(plan)]["xcd_chunk"] == 4


def test_committed_cache_covers_the_tuned_workloads():
    cache = AT.load_cache(AT.CACHE)
    keys = {_strip_device(k) for k in cache}
    for key in ("gen|Roe|4|8x150x150", "cart|HLLL|2|5x500x500", "cart|Roe|2|1x256x256", "cart|Roe|4|1x100x100",
                "mixed|HLLL/prim|2|4x50x50"):
        assert key in keys, key
    for k, e in cache.items():
        assert k.endswith("|gfx950")
        assert 0 <= e["variant"] <= 8 and e["rows_per_item"] >= 1 and e["waves_per_simd"] in (2, 3, 4)
        if e["variant"] in (6, 7):
            assert 1 <= e["rows_b"] <= YU_MAX_ROWS
        if e["variant"] == 8:
            assert 1 <= e["rows_b"] <= FLAT_MAX_WAVES - 1
        assert "/q" not in k or e["variant"] in (1, 5, 6)   # Gauss faces run the split stages only


def test_autotune_skips_refused_shapes_and_always_restores_the_plan(tmp_path, monkeypatch):
    """A shape the library refuses (variant 8 on blocks wider than its LDS ranges) is skipped and
    recorded; the trials' state changes are undone even when a trial fails midway."""
    plan = fake_plan([True], [(2100, 8)])
    plan.U = torch.arange(8, dtype=torch.float64)
    state = {"launch": None, "resets": []}
    plan.status = lambda: (None, 3, 0.5, None)
    plan.reset = lambda t0, step0: state["resets"].append((float(t0), step0))

    def set_launch(r, g, w, variant=0, rows_b=0, xcd_chunk=0):
        if variant == 8:
            raise ValueError("block too wide for the flat y-face ranges")
        state["launch"] = (variant, r, g, w, rows_b)

    plan.set_launch = set_launch

    def timing(p, steps):
        p.U.add_(1.0)          # a trial advances the state
        return 1.0 - 0.5 * (state["launch"][0] == 6)

    monkeypatch.setattr(AT, "_time_steps", timing)
    best = AT.autotune(plan, steps=1, path=tmp_path / "tuning.json")
    assert best["variant"] == 6
    assert best["skipped"] and all(c["variant"] == 8 for c in best["skipped"])
    assert torch.equal(plan.U, torch.arange(8, dtype=torch.float64))
    assert state["resets"][-1] == (0.5, 3)

    def failing(p, steps):
        p.U.add_(1.0)
        raise RuntimeError("trial failed")

    monkeypatch.setattr(AT, "_time_steps", failing)
    with pytest.raises(RuntimeError):
        AT.autotune(plan, steps=1, path=None)
    assert torch.equal(plan.U, torch.arange(8, dtype=torch.float64))
    assert state["resets"][-1] == (0.5, 3)


def test_candidates_include_persistent_grids():
    cands = AT.candidates()
    assert {c["grid_waves"] for c in cands if c["variant"] == 0} == {0, 1024, 2048, 3072}
    assert {c["variant"] for c in cands if c["grid_waves"] > 0} == {0, 6, 7, 8}
