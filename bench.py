#!/usr/bin/env python3
"""Benchmark of the MUSCL residual + explicit-RK step on one MI355X.

Default workload: the multiblock explosion (2x4 blocks of 150x150 on [0,10]x[0,20], Roe,
Venkatakrishnan, Green-Gauss, RK4 at CFL 0.8, reflecting walls, FP64).  A "step" is one full time
step (CFL dt, four RK stages over all blocks).  ``value`` is FP64 cell-stage updates per second of
the whole job with the state resident in HBM.  ``roofline`` prices the stage kernel's algorithmic
bytes (compulsory FP64 state traffic per cell-stage) against the 8 TB/s HBM peak, with the kernel
duration measured by HIP events on the launch stream.  ``cpu_baseline`` times the NumPy oracle on
the host on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "FP64 cell-stage updates/s + % HBM peak; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
PROFILES = ROOT / "profiles"


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


def build_workload(name: str, args):
    from pyhype_amd import Euler2D, Euler2DBatch
    from pyhype_amd import workloads as W

    tune = {"rows_per_item": args.rows, "grid_waves": args.grid}
    if name == "explosion_batch":
        cfgs, meshes, params = W.explosion_batch(seed=args.seed)
        solver = Euler2DBatch(cfgs, meshes, **tune)
        desc = {"workload": name, "cases": len(cfgs), "seed": args.seed,
                "sides": [c["n"] for c in params["cases"]], "scheme": "Roe/Venkatakrishnan/GG, RK4, CFL 0.8"}
        return solver, cfgs[0], meshes, desc
    fn = getattr(W, name)
    cfg, mesh = fn()
    solver = Euler2D(cfg, mesh, **tune)
    nb = len(solver.plan.blocks)
    desc = {"workload": name, "blocks": nb, "nx": cfg.nx, "ny": cfg.ny,
            "scheme": f"{cfg.fvm_flux_function_type}/Venkatakrishnan/GG, {cfg.time_integrator}, CFL {cfg.CFL}",
            "geometry": "cartesian" if all(e.cartesian for e in solver.plan.blocks) else "general"}
    return solver, cfg, mesh, desc


def cpu_baseline(name: str, seconds: float):
    """Oracle (NumPy restatement, single thread) on a bounded sample of the same workload."""
    import numpy as np

    from oracle import pyhype_oracle as O
    from pyhype_amd import Euler2D
    from pyhype_amd import workloads as W
    from pyhype_amd.solvers import mesh_dict
    from pyhype_amd.time_marching import butcher_table

    if name == "explosion_batch":
        cfgs, meshes, _ = W.explosion_batch(seed=0)
        k = int(np.argmax([c.nx for c in cfgs]))
        cfg, mesh = cfgs[k], meshes[k]
        sample = f"case {k} ({cfg.nx}x{cfg.ny}) of the batch"
    else:
        cfg, mesh = getattr(W, name)()
        sample = f"all {len(mesh_dict(mesh))} block(s) of {cfg.nx}x{cfg.ny}"
    from pyhype_amd.mesh import BlockGeometry
    from pyhype_amd.solvers import HostBlock

    md = mesh_dict(mesh)
    sim = O.Simulation(md, cfg.nx, cfg.ny, flux=cfg.fvm_flux_function_type,
                       integrator=butcher_table(cfg.time_integrator, cfg), CFL=cfg.CFL, t_final=cfg.t_final,
                       a_inf=cfg.fluid.far_field.a)

    class _E:  # minimal entry for HostBlock
        def __init__(self, n, g):
            self.num, self.geom, self.nx, self.ny = n, g, g.nx, g.ny

    states = {}
    for n, d in md.items():
        g = BlockGeometry({k2: d[k2] for k2 in ("SW", "SE", "NW", "NE")}, cfg.nx, cfg.ny)
        hb = HostBlock(cfg, _E(n, g))
        cfg.initial_condition.apply_to_block(hb)
        states[n] = hb.state.data
    sim.set_state(states)
    cells = sum(cfg.nx * cfg.ny for _ in md)
    stages = len(butcher_table(cfg.time_integrator, cfg))
    t0 = time.perf_counter()
    steps = 0
    while True:
        sim.step()
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 50:
            break
    return {"value": cells * stages * steps / el, "unit": "cell-stage updates/s", "cores": 1, "kind": "port",
            "sample": f"NumPy oracle, {sample}, {steps} step(s) from the initial state, {el:.1f} s"}


def adjoint_ratio(n: int = 256, steps: int = 100, every: int = 10, reps: int = 3):
    """Adjoint/forward time ratio on the discrete-adjoint workload (256x256 smoothed explosion,
    Roe, SSP-RK2, CFL 0.4, 100 frozen-dt steps, checkpoints every 10): T(full reverse sweep,
    checkpoint recompute included) / T(forward of the same 100 steps), HIP events on the launch
    stream, best of ``reps``."""
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
            "adjoint_ms": round(best_b, 4), "ratio": round(best_b / best_f, 4), "J": J}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="explosion_multiblock")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--rows", type=int, default=0, help="rows per work item (0 = tuned/default)")
    ap.add_argument("--grid", type=int, default=0, help="persistent grid in waves (0 = one per item)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tune", action="store_true", help="autotune the launch shape before timing")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward ratio measurement")
    ap.add_argument("--pmc-calib", action="store_true",
                    help="also run a 256 MiB device copy three times (PMC byte-counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        log("bench.py runs on one GPU only: this build has no multi-GPU path (--gpus must be 1)")
        sys.exit(2)

    import torch

    from pyhype_amd import workloads as W
    from pyhype_amd.time_marching import butcher_table

    default_steps = W.WORKLOADS[args.workload][1]
    K = args.steps if args.steps is not None else default_steps
    Wm = args.warmup
    solver, cfg, mesh, desc = build_workload(args.workload, args)
    plan = solver.plan
    table = butcher_table(cfg.time_integrator, cfg)
    S = len(table)
    cells = plan.cells()
    solver.apply_initial_condition()
    tuned = "cache" if getattr(solver, "tuned", False) else "default"
    if args.tune:
        solver.run_steps(2)
        res = solver.autotune(steps=8)
        tuned = "autotuned"
        log(f"autotune best {res['rows_per_item']} rows, {res['waves_per_simd']} waves/SIMD: "
            f"{res['ms_per_step']:.4f} ms/step")
    elif args.rows:
        tuned = "manual"
    log(f"workload {desc}; cells {cells}; launch {plan.launch_info()} ({tuned})")

    use_graph = not args.no_graph
    strict = getattr(solver, "strict", True)
    plan.run(Wm, use_graph=use_graph, check=strict)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record()
    plan.run(K, use_graph=use_graph, check=False)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = ev0.elapsed_time(ev1)
    failed = plan.failed_cases()
    if failed and strict:
        plan.check_realizable()
    # a case that became unrealizable is frozen (the reference aborts that run): count live work only
    live = cells - sum(e.nx * e.ny for e in plan.blocks if e.case in failed)
    value = live * S * K / wall
    if failed:
        desc["unrealizable_cases"] = failed
        desc["live_cells"] = live

    # stage-kernel duration with HIP events on the launch stream: the stage programs of `reps`
    # steps launched back to back between two events (the average includes the short launch gaps)
    reps = max(3, min(50, int(2000 // max(1, K))))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for r in range(reps):
        for s in range(S):
            plan.stage(s)
    e1.record()
    torch.cuda.synchronize()
    stage_ms = e0.elapsed_time(e1) / (reps * S)
    b_cell_step = alg_bytes_per_cell_step(table)
    # the stage launches above skip the blocks of frozen cases: credit the live cells only
    bytes_per_launch = live * b_cell_step / S
    achieved = bytes_per_launch / (stage_ms * 1e-3) / 1e9
    traffic = None
    tfile = PROFILES / f"traffic_{args.workload}.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    variant = plan.launch_info().get("variant", 0)
    kernel = {1: "k_recon_x + k_yface + k_update (split stage: one RK stage over all blocks)",
              2: "k_stage_wg (shared-halo stage: one RK stage over all blocks)",
              3: "k_ew_recon + k_ew_face + k_ew_update (element-wise stage: one RK stage over all blocks)",
              4: "k_stage_pair (paired stage: one RK stage over all blocks)",
              5: "k_recon_x + k_yface_update (split stage with fused y-faces and update: one RK stage over all "
                 "blocks)",
              6: "k_recon_x + k_yface_update_wg (split stage, y-faces and update with one face per wave: one RK "
                 "stage over all blocks)",
              7: "k_recon_x_lin + k_yface_update_wg (split stage, row-major reconstruction, one y-face per wave: "
                 "one RK stage over all blocks)",
              8: "k_recon_x_lin + k_yface_update_flat (split stage, row-major reconstruction, y-faces and update "
                 "over flat cell ranges: one RK stage over all blocks)"}.get(
                  variant, "k_stage (one RK stage over all blocks)")
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": kernel, "stage_ms": round(stage_ms, 5),
                "alg_bytes_per_cell_stage": b_cell_step / S}

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(args.workload, args.cpu_seconds)
        except Exception as exc:  # the baseline never blocks the GPU figure
            log(f"cpu baseline failed: {exc!r}")
    desc.update({"steps_timed": K, "graph": use_graph, "launch": plan.launch_info(), "launch_source": tuned})
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
        desc["pmc_calibration"] = "3 x copy of 256 MiB (268435456 B read + written per launch)"
    out = {
        "metric": METRIC, "value": value, "unit": "cell-stage updates/s", "n_gpus": 1, "steps": K, "warmup": Wm,
        "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic", "config": desc, "roofline": roofline, "cpu_baseline": cpu,
        "hbm_frac_of_value": value * (b_cell_step / S) / (HBM_PEAK_GBS * 1e9), "event_ms": ev_ms,
        "adjoint": adjoint,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
