#!/usr/bin/env python3
"""Benchmark of the MUSCL residual + explicit-RK step on one MI355X.

Default workload: the 64 batched uneven explosion cases (seed 0; sides 64..512, 7.79 M cells;
Roe, Venkatakrishnan, Green-Gauss, RK4 at CFL 0.8, reflecting walls, FP64, per-case dt), the
largest single-GPU configuration and the one whose working set (about 1 GB per stage) does not
fit the 256 MB last-level cache.  A "step" is one full time step of every case (CFL dt, four RK
stages).  ``value`` is FP64 cell-stage updates per second of the whole job with the state
resident in HBM; a case the scheme cannot advance becomes unrealizable and is frozen (the
reference aborts that run), and only the cells of running cases are counted.

``roofline`` prices the stage kernel's algorithmic bytes (compulsory FP64 state traffic per
cell-stage of the live cells) against the 8 TB/s HBM peak, with the kernel duration measured by
HIP events on the launch stream; ``roofline.valu`` prices the kernel's FP64 instructions (per live
cell-stage, from the SQ counters kept under profiles/) against the FP64 issue rate.
``cpu_baseline`` times the NumPy oracle on the host: one core, and all cores (one process per
core, each on its own case).  ``extra`` carries the other workloads (8x150^2 multiblock
explosion, 5x500^2 HLLL oblique shock, 100^2 single block, the 256^2 SSP-RK2 forward of the
adjoint problem) and ``adjoint`` the discrete-adjoint costs on that 256^2 problem.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "FP64 cell-stage updates/s + % HBM peak; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
# FP64 VALU issue rate of MI355X: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz (78.6 TFLOP/s counting FMA twice)
F64_ISSUE_PEAK = 256 * 4 * 16 * 2.4e9
PROFILES = ROOT / "profiles"
EXTRA = ("explosion_multiblock", "oblique_shock", "explosion_single", "explosion_adjoint")
MAX_CPU_CORES = 16   # the host-core share of one GPU


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alg_bytes_per_cell_step(table) -> float:
    """Compulsory FP64 state traffic per cell and step of the accumulator-form RK program: per stage
    one read of every distinct input (stage state, U0, accumulators) and one write of every output."""
    from pyhype_amd.backend import stage_programs

    total = 0
    for prog in stage_programs(table):
        reads = {prog["in"]} | {o["src"] for o in prog["outs"] if o["a"] != 0.0 or o["src"] != o["dst"]}
        writes = {o["dst"] for o in prog["outs"]}
        total += 32 * (len(reads) + len(writes))
    return float(total)


KERNELS = {1: "k_recon_x + k_yface + k_update (split stage: one RK stage over all blocks)",
           2: "k_stage_wg (shared-halo stage: one RK stage over all blocks)",
           3: "k_ew_recon + k_ew_face + k_ew_update (element-wise stage: one RK stage over all blocks)",
           4: "k_stage_pair (paired stage: one RK stage over all blocks)",
           5: "k_recon_x + k_yface_update (split stage with fused y-faces and update: one RK stage over all blocks)",
           6: "k_recon_x + k_yface_update_wg (split stage, y-faces and update with one face per wave: one RK stage "
              "over all blocks)",
           7: "k_recon_x_lin + k_yface_update_wg (split stage, row-major reconstruction, one y-face per wave: one RK "
              "stage over all blocks)",
           8: "k_recon_x_lin + k_yface_update_flat (split stage, row-major reconstruction, y-faces and update over "
              "flat cell ranges: one RK stage over all blocks)"}


def build_workload(name: str, args):
    from pyhype_amd import Euler2D, Euler2DBatch
    from pyhype_amd import workloads as W

    tune = {"rows_per_item": args.rows, "grid_waves": args.grid}
    if name == "explosion_batch":
        cfgs, meshes, params = W.explosion_batch(seed=args.seed)
        solver = Euler2DBatch(cfgs, meshes, **tune)
        desc = {"workload": name, "cases": len(cfgs), "seed": args.seed,
                "sides": [c["n"] for c in params["cases"]], "scheme": "Roe/Venkatakrishnan/GG, RK4, CFL 0.8"}
        return solver, cfgs[0], desc
    cfg, mesh = getattr(W, name)()
    solver = Euler2D(cfg, mesh, **tune)
    desc = {"workload": name, "blocks": len(solver.plan.blocks), "nx": cfg.nx, "ny": cfg.ny,
            "scheme": f"{cfg.fvm_flux_function_type}/{cfg.fvm_slope_limiter_type}/GG, {cfg.time_integrator}, "
                      f"CFL {cfg.CFL}",
            "geometry": "cartesian" if all(e.cartesian for e in solver.plan.blocks) else "general"}
    return solver, cfg, desc


def measure(name: str, args, steps: int, warmup: int, tune: bool = False):
    """Timed steps of one workload (hipGraph replay, wall clock between synchronisations), then the
    stage kernel's duration from HIP events around back-to-back stage launches of the final state."""
    import torch

    from pyhype_amd.time_marching import butcher_table

    solver, cfg, desc = build_workload(name, args)
    plan = solver.plan
    table = butcher_table(cfg.time_integrator, cfg)
    S = len(table)
    cells = plan.cells()
    solver.apply_initial_condition()
    source = "cache" if getattr(solver, "tuned", False) else "default"
    if tune:
        solver.run_steps(2)
        res = solver.autotune(steps=8)
        source = "autotuned"
        log(f"autotune {name}: best {res['variant']}/{res['rows_per_item']}/{res['rows_b']}/{res['grid_waves']}: "
            f"{res['ms_per_step']:.4f} ms/step")
        solver.apply_initial_condition()
    elif args.rows:
        source = "manual"
    log(f"workload {desc}; cells {cells}; launch {plan.launch_info()} ({source})")
    use_graph = not args.no_graph
    strict = getattr(solver, "strict", True)
    plan.run(warmup, use_graph=use_graph, check=strict)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    plan.run(steps, use_graph=use_graph, check=False)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    failed = plan.failed_cases()
    if failed and strict:
        plan.check_realizable()
    # a case that became unrealizable is frozen (the reference aborts that run): count live work only
    live = cells - sum(e.nx * e.ny for e in plan.blocks if e.case in failed)
    if failed:
        desc["unrealizable_cases"] = {int(c): int(s) for c, s in failed.items()}
        desc["live_cells"] = live
    # stage-kernel duration with HIP events on the launch stream (torch's current stream, which the
    # plan launches on): the stage programs of `reps` steps back to back on the final state
    reps = max(3, min(50, int(2000 // max(1, steps))))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        for s in range(S):
            plan.stage(s)
    e1.record()
    torch.cuda.synchronize()
    stage_ms = e0.elapsed_time(e1) / (reps * S)
    b_cell_step = alg_bytes_per_cell_step(table)
    desc.update({"steps_timed": steps, "graph": use_graph, "launch": plan.launch_info(), "launch_source": source})
    out = {"value": live * S * steps / wall, "ms_per_step": wall * 1e3 / steps, "event_ms": ev0.elapsed_time(ev1),
           "stage_ms": stage_ms, "stage_launches_timed": reps * S, "cells": cells, "live_cells": live, "stages": S,
           "alg_bytes_per_cell_stage": b_cell_step / S, "variant": plan.launch_info().get("variant", 0),
           "config": desc}
    plan.close()
    return out


def roofline(m: dict, workload: str):
    """HBM roofline of the stage kernel plus its FP64-issue ceiling (from the committed counters)."""
    bytes_per_launch = m["live_cells"] * m["alg_bytes_per_cell_stage"]
    achieved = bytes_per_launch / (m["stage_ms"] * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
           "kernel": KERNELS.get(m["variant"], "k_stage (fused: one RK stage over all blocks)"),
           "stage_ms": round(m["stage_ms"], 5), "alg_bytes_per_cell_stage": m["alg_bytes_per_cell_stage"],
           "alg_bytes_per_launch": bytes_per_launch, "live_cells": m["live_cells"]}
    tfile = PROFILES / f"traffic_{workload}.json"
    if tfile.exists():
        try:
            t = json.loads(tfile.read_text())
            # HBM bytes per live cell-stage of the profiled launches, times this run's live cells
            out["traffic"] = int(round(t["hbm_bytes_per_cell_stage"] * m["live_cells"]))
            out["traffic_source"] = f"profiles/{tfile.name}"
        except (OSError, ValueError, KeyError):
            pass
    sfile = PROFILES / f"sq_{workload}.json"
    if sfile.exists():
        try:
            q = json.loads(sfile.read_text())
            per_cell = q["f64_lane_instr_per_cell_stage"]
            rate = per_cell * m["live_cells"] / (m["stage_ms"] * 1e-3)
            out["valu"] = {"bound": "fp64-issue", "achieved": round(rate / 1e12, 3), "peak": round(F64_ISSUE_PEAK / 1e12, 3),
                           "unit": "T FP64 lane-instructions/s", "frac": round(rate / F64_ISSUE_PEAK, 4),
                           "f64_lane_instr_per_cell_stage": per_cell,
                           "valu_lane_instr_per_cell_stage": q.get("valu_lane_instr_per_cell_stage"),
                           "source": f"profiles/{sfile.name}"}
        except (OSError, ValueError, KeyError):
            pass
    return out


# ---------------------------------------------------------------------------------------------
# CPU baseline: the NumPy oracle on the host cores
# ---------------------------------------------------------------------------------------------
def _oracle_case(workload: str, case: int, seed: int):
    """Oracle simulation of one case of a workload at its initial state, and its cells x stages."""
    from oracle import pyhype_oracle as O
    from pyhype_amd import workloads as W
    from pyhype_amd.solvers import host_initial_states, mesh_dict
    from pyhype_amd.time_marching import butcher_table

    if workload == "explosion_batch":
        cfgs, meshes, _ = W.explosion_batch(seed=seed)
        cfg, mesh = cfgs[case], meshes[case]
    else:
        cfg, mesh = getattr(W, workload)()
    md = mesh_dict(mesh)
    table = butcher_table(cfg.time_integrator, cfg)
    sim = O.Simulation(md, cfg.nx, cfg.ny, flux=cfg.fvm_flux_function_type, integrator=table, CFL=cfg.CFL,
                       t_final=cfg.t_final, a_inf=cfg.fluid.far_field.a, limiter=cfg.fvm_slope_limiter_type)
    sim.set_state(host_initial_states(cfg, mesh))
    return sim, cfg.nx * cfg.ny * len(md) * len(table), f"{len(md)} block(s) of {cfg.nx}x{cfg.ny}"


def _oracle_timed(job):
    """Runs oracle steps of one case for about `seconds`: (cell-stages done, seconds, steps, label)."""
    workload, case, seed, seconds = job
    sim, per_step, label = _oracle_case(workload, case, seed)
    t0 = time.perf_counter()
    steps = 0
    while True:
        try:
            sim.step()
        except FloatingPointError:  # the case failed: the steps done so far are the sample
            break
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 200:
            break
    el = time.perf_counter() - t0
    return per_step * steps, el, steps, label


def _pool_init():
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ[k] = "1"


def cpu_baseline(workload: str, seed: int, seconds: float):
    """NumPy oracle (single-threaded per process) on a bounded sample of the same workload: one core
    (the largest case of the batch, or the whole mesh), then all cores (one process per core, each
    running its own case of the batch -- or its own copy of the mesh -- for the same time)."""
    import numpy as np

    from pyhype_amd import workloads as W

    if workload == "explosion_batch":
        sides = [c.nx for c in W.explosion_batch(seed=seed)[0]]
        big = int(np.argmax(sides))
        order = [int(c) for c in np.argsort(sides)[::-1]]
    else:
        big, order = 0, [0]
    done, el, steps, label = _oracle_timed((workload, big, seed, seconds))
    one = {"value": done / el, "unit": "cell-stage updates/s", "cores": 1, "kind": "port",
           "sample": f"NumPy oracle, one process: case {big} ({label}), {steps} step(s) from the initial state, "
                     f"{el:.1f} s"}
    cores = max(1, min(MAX_CPU_CORES, len(os.sched_getaffinity(0))))
    jobs = [(workload, order[i % len(order)], seed, seconds) for i in range(cores)]
    ctx = mp.get_context("spawn")   # the parent holds a GPU context: children start fresh
    t0 = time.perf_counter()
    with ctx.Pool(cores, initializer=_pool_init) as pool:
        res = pool.map(_oracle_timed, jobs, chunksize=1)
    wall = time.perf_counter() - t0
    total = sum(r[0] / r[1] for r in res)
    allc = {"value": total, "unit": "cell-stage updates/s", "cores": cores, "kind": "port",
            "sample": f"NumPy oracle, {cores} processes at once, each on its own "
                      + ("case of the batch (the largest first)" if workload == "explosion_batch"
                         else "copy of the mesh") + f" for about {seconds:.0f} s from the initial state "
                      f"({sum(r[2] for r in res)} steps in all, {wall:.1f} s with start-up); value = sum of the "
                      "per-process rates"}
    return dict(allc, one_core=one)


# ---------------------------------------------------------------------------------------------
# discrete adjoint
# ---------------------------------------------------------------------------------------------
def adjoint_ratio(n: int = 256, steps: int = 100, every: int = 10, reps: int = 3):
    """Adjoint/forward time ratio on the discrete-adjoint workload (256x256 smoothed explosion,
    Roe, SSP-RK2, CFL 0.4, 100 frozen-dt steps, checkpoints every 10): T(full reverse sweep,
    checkpoint recompute included) / T(forward of the same 100 steps), HIP events on the launch
    stream, best of ``reps``; also the whole gradient (checkpointed forward + sweep) / forward."""
    import torch

    from pyhype_amd import Euler2D
    from pyhype_amd import workloads as W
    from pyhype_amd.adjoint import AdjointProblem

    cfg, mesh, target, params = W.adjoint_problem(n, seed=0)
    prob = AdjointProblem(Euler2D(cfg, mesh), target, nsteps=steps, checkpoint_every=every)
    P = prob.plan
    J, g = prob.gradient()  # warm-up (allocates the adjoint work arrays)
    lam = torch.zeros_like(P.U)
    lam.view(-1, 4)[:, 0] = 1.0
    best_f = best_b = best_fc = float("inf")
    for _ in range(reps):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        P.U.copy_(prob.U0)
        torch.cuda.synchronize()
        ev[0].record()
        P.run_frozen(prob.dt, check=False)
        ev[1].record()
        ev[2].record()
        prob.traj.forward(prob.U0)
        ev[3].record()
        ev[4].record()
        prob.traj.backward(lam)
        ev[5].record()
        torch.cuda.synchronize()
        best_f = min(best_f, ev[0].elapsed_time(ev[1]))
        best_fc = min(best_fc, ev[2].elapsed_time(ev[3]))
        best_b = min(best_b, ev[4].elapsed_time(ev[5]))
    P.close()
    return {"workload": f"adjoint_problem({n}), {steps} SSP-RK2 steps, checkpoints every {every}",
            "forward_ms": round(best_f, 4), "forward_with_checkpoints_ms": round(best_fc, 4),
            "adjoint_ms": round(best_b, 4), "ratio": round(best_b / best_f, 4),
            "gradient_ms": round(best_fc + best_b, 4), "gradient_ratio": round((best_fc + best_b) / best_f, 4),
            "J": J}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="explosion_batch")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--rows", type=int, default=0, help="rows per work item (0 = tuned/default)")
    ap.add_argument("--grid", type=int, default=0, help="persistent grid in waves (0 = one per item)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tune", action="store_true", help="autotune the launch shape before timing")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward ratio measurement")
    ap.add_argument("--no-extra", action="store_true", help="skip the other workloads")
    ap.add_argument("--pmc-calib", action="store_true",
                    help="also run a 256 MiB device copy three times (PMC byte-counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        log("bench.py runs on one GPU only: this build has no multi-GPU path (--gpus must be 1)")
        sys.exit(2)

    import torch

    from pyhype_amd import workloads as W

    K = args.steps if args.steps is not None else W.WORKLOADS[args.workload][1]
    m = measure(args.workload, args, K, args.warmup, tune=args.tune)
    roof = roofline(m, args.workload)
    log(f"{args.workload}: {m['value']:.4g} cell-stage/s, stage {m['stage_ms'] * 1e3:.1f} us, "
        f"HBM frac {roof['frac']:.4f}")

    extra = {}
    if not args.no_extra:
        for w in EXTRA:
            if w == args.workload:
                continue
            try:
                e = measure(w, args, W.WORKLOADS[w][1], 2)
                r = roofline(e, w)
                extra[w] = {"value": e["value"], "ms_per_step": e["ms_per_step"], "steps": W.WORKLOADS[w][1],
                            "stage_ms": round(e["stage_ms"], 5), "hbm_frac": r["frac"], "achieved_gbs": r["achieved"],
                            "kernel": r["kernel"], "launch": e["config"]["launch"]}
                if "valu" in r:
                    extra[w]["valu_frac"] = r["valu"]["frac"]
                log(f"{w}: {e['value']:.4g} cell-stage/s, stage {e['stage_ms'] * 1e3:.1f} us, HBM frac {r['frac']:.4f}")
            except Exception as exc:  # reported, never blocks the headline figure
                log(f"{w} failed: {exc!r}")
                extra[w] = {"error": repr(exc)}

    adjoint = None
    if not args.no_adjoint:
        try:
            adjoint = adjoint_ratio()
        except Exception as exc:  # reported, never blocks the forward figure
            log(f"adjoint measurement failed: {exc!r}")
            adjoint = {"error": repr(exc)}
    if args.pmc_calib:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        m["config"]["pmc_calibration"] = "3 x copy of 256 MiB (268435456 B read + written per launch)"

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(args.workload, args.seed, args.cpu_seconds)
        except Exception as exc:  # the baseline never blocks the GPU figure
            log(f"cpu baseline failed: {exc!r}")
    out = {
        "metric": METRIC, "value": m["value"], "unit": "cell-stage updates/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": m["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic", "config": m["config"], "roofline": roof,
        "cpu_baseline": cpu,
        "hbm_frac_of_value": m["value"] * m["alg_bytes_per_cell_stage"] / (HBM_PEAK_GBS * 1e9),
        "event_ms": m["event_ms"], "extra": extra, "adjoint": adjoint,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
