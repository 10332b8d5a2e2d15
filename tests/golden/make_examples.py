"""Writes the digests of the reference's example runs (tests/golden/example_<name>.npz): each example's
configuration, mesh and initial condition run to its final time with the oracle
(oracle/pyhype_oracle.py), through the step loop of Euler2D._solve (solvers/Euler2D.py:195-210):
``while t < t_final: dt = get_dt(); integrate(dt); realizability; t += dt``, dt clipped to
t_final - t (solvers/base.py:132-136), t_final scaled by a_inf (solvers/base.py:73).

* ``explosion``: examples/explosion (config.py, mesh.py): one 40x40 block on [0,10]x[0,20], Roe,
  RK4, CFL 0.7, t_final 0.07;
* ``explosion_multi``: examples/explosion_multi: 2x4 blocks of 150x150 on [0,10]x[0,20]
  (RectagularMeshGenerator), Roe, RK4, CFL 0.7, t_final 0.07;
* ``implosion``: examples/implosion: one 40x40 block on [0,10]^2, Roe, RK2, CFL 0.4, t_final 0.1;
* ``shockbox``: examples/shockbox: one 50x50 block on [0,10]^2 numbered 1 (the example's mesh
  dictionary has the single key 1), Roe, RK2, CFL 0.4, t_final 2.0;
* ``dmr``: examples/dmr: 4x1 blocks of 50x50 on the 30-degree ramp, HLLL, RK2, PrimitiveState
  reconstruction, CFL 0.4, t_final 0.25.

A run that becomes unrealizable stops at its failing step, as the reference's _realizability_check
aborts it (solvers/Euler2D.py:144-152); the file then records that step and the state it produced.

Each file keeps the number of steps, the final time, every step's dt, the sha256 of every block's
initial and final state bytes (NaNs canonical), per-block maxima and sums and 64 seeded sample
cells per block (the same digest as tests/golden/make_fullsize.py).

Run from the repository root (one process per example; ``explosion_multi`` takes about an hour, the
others under a minute each):
    python tests/golden/make_examples.py [name ...]
"""
from __future__ import annotations

import sys
import time
from multiprocessing import Pool
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from tests.golden.make_fullsize import sha, summarize  # noqa: E402

HERE = Path(__file__).resolve().parent
EXAMPLES = ("explosion", "explosion_multi", "implosion", "shockbox", "dmr")
MAX_STEPS = {}   # example -> step cap (none needed: every example ends within 2 000 steps)


def example(name):
    """(config, mesh, step cap or None) of one of the reference's examples."""
    from pyhype_amd import workloads as W
    from pyhype_amd.initial_conditions import ExplosionInitialCondition

    if name == "explosion":
        cfg = W._config(40, 40, ic=ExplosionInitialCondition(), integrator="RK4", cfl=0.7)
        mesh = W.single_block_mesh(east=10.0, north=20.0)
    elif name == "explosion_multi":
        cfg, mesh = W.explosion_multiblock(150)
        cfg.CFL = 0.7
    elif name == "implosion":
        cfg, mesh = W.implosion(40)
    elif name == "shockbox":
        cfg, mesh = W.shockbox(50)
        blk = dict(mesh[0])
        blk["nBLK"] = 1
        mesh = {1: blk}
    elif name == "dmr":
        cfg, mesh = W.double_mach_reflection(50)
    else:
        raise KeyError(name)
    cfg.t_final = {"explosion": 0.07, "explosion_multi": 0.07, "implosion": 0.1, "shockbox": 2.0,
                   "dmr": 0.25}[name]
    return cfg, mesh, MAX_STEPS.get(name)


def run(name):
    """Oracle run of one example: digest arrays."""
    from pyhype_amd.solvers import host_initial_states
    from tests.helpers import oracle_for

    cfg, mesh, cap = example(name)
    U0 = host_initial_states(cfg, mesh)
    sim = oracle_for(cfg, mesh, U0)
    failed = -1
    dts = []
    t0 = time.perf_counter()
    while sim.t < sim.t_final and (cap is None or sim.nsteps < cap):
        dt = sim.get_dt()
        try:
            sim.step(dt)
        except FloatingPointError:
            failed = sim.nsteps + 1
            dts.append(dt)
            break
        dts.append(dt)
    el = time.perf_counter() - t0
    print(f"{name}: {len(sim.blocks)} block(s) of {cfg.nx}x{cfg.ny}, {len(dts)} steps, t = {sim.t!r} of "
          f"{sim.t_final!r}, failed {failed}, {el:.1f} s", flush=True)
    d = summarize(sim.states())
    arrays = {"steps": np.array(len(dts)), "cap": np.array(-1 if cap is None else cap), "failed": np.array(failed),
              "t": np.array(sim.t), "t_final": np.array(sim.t_final), "dts": np.array(dts),
              "blocks": np.array(list(sim.blocks)), "sha0": np.array([sha(U0[n]) for n in sim.blocks]),
              "sha": np.array(d["sha"]), "absmax": np.array(d["absmax"]), "sum": np.array(d["sum"]),
              "samples": np.array(d["samples"])}
    np.savez_compressed(HERE / f"example_{name}.npz", **arrays)
    return name, el


def main():
    names = sys.argv[1:] or list(EXAMPLES)
    with Pool(min(len(names), 5)) as pool:
        for name, el in pool.imap_unordered(run, names):
            print(f"{name} written ({el:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
