"""Writes the regression fixtures of the oracle (tests/golden/*.npz): for a few small seeded cases,
the initial state, the residual of that state and the state after a few steps, as computed by
oracle/pyhype_oracle.py.  The reference holds no golden vectors for this path (it has no tests or
data files), so these fixtures pin the oracle against itself: a later edit of the oracle that
changes any bit of these results fails tests/test_golden.py, and the GPU tests compare the HIP
path with the same committed arrays.

Run from the repository root:  python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import pyhype_oracle as O  # noqa: E402

HERE = Path(__file__).resolve().parent


def _blk(n, sw, se, nw, ne, nb, bc):
    d = {"nBLK": n, "SW": sw, "SE": se, "NW": nw, "NE": ne}
    for s in O.DIRS:
        d["Neighbor" + s] = nb.get(s)
        d["BCType" + s] = bc.get(s)
    return d


def cases():
    """name -> (mesh dict, nx, ny, Simulation kwargs, steps)."""
    walls = {s: "Reflection" for s in O.DIRS}
    one = {0: _blk(0, [0.0, 0.0], [10.0, 0.0], [0.0, 10.0], [10.0, 10.0], {}, walls)}
    two = {0: _blk(0, [0.0, 0.0], [5.0, 0.8], [0.3, 10.0], [5.6, 9.1], {"E": 1},
                   {"W": "OutletDirichlet", "N": "Reflection", "S": "Reflection"}),
           1: _blk(1, [5.0, 0.8], [10.0, 1.9], [5.6, 9.1], [10.4, 8.0], {"W": 0},
                   {"E": "Reflection", "N": "Reflection", "S": "Reflection"})}
    return {
        "roe_rk4_walls_12": (one, 12, 12, dict(flux="Roe", integrator="RK4", CFL=0.8), 3),
        "hlll_ssprk2_skewed_10": (two, 10, 10, dict(flux="HLLL", integrator="SSPRK2", CFL=0.4), 3),
        "roe_vanleer_prim_gauss2_skewed_10": (two, 10, 10, dict(flux="Roe", integrator="RK4", CFL=0.6,
                                                                limiter="VanLeer", reconstruction="Primitive",
                                                                quadrature_points=2), 2),
        "roe_barth_gauss3_12": (one, 12, 12, dict(flux="Roe", integrator="SSPRK2", CFL=0.5,
                                                  limiter="BarthJespersen", quadrature_points=3), 2),
    }


def initial_state(mesh, nx, ny, seed):
    """A smooth, seeded, non-symmetric state: explosion-like pressure bump plus a shear flow."""
    rng = np.random.default_rng(seed)
    out = {}
    for n, d in mesh.items():
        g = O.block_geometry({k: d[k] for k in ("SW", "SE", "NW", "NE")}, nx, ny)
        x, y = g.xc, g.yc
        cx, cy = 5.0 + rng.uniform(-1, 1), 5.0 + rng.uniform(-1, 1)
        r2 = (x - cx) ** 2 + (y - cy) ** 2
        rho = 1.0 + 0.5 * np.exp(-r2 / 4.0)
        u = 0.2 * np.sin(0.3 * y) + 0.05 * rng.standard_normal(x.shape)
        v = 0.1 * np.cos(0.2 * x)
        p = 1.0 + 2.0 * np.exp(-r2 / 3.0)
        W = np.stack([rho, u, v, p], axis=-1)
        out[n] = O.prim_to_cons(W)
    return out


def compute(name):
    mesh, nx, ny, kw, steps = cases()[name]
    sim = O.Simulation(mesh, nx, ny, **kw)
    U0 = initial_state(mesh, nx, ny, seed=sum(map(ord, name)))
    sim.set_state(U0)
    R = {n: O.residual(b, sim.flux_fn) for n, b in sim.blocks.items()}
    sim.run(steps)
    return U0, R, sim.states(), np.array(sim.dts)


def main():
    for name in cases():
        U0, R, U, dts = compute(name)
        arrays = {"dts": dts}
        for n in U0:
            arrays[f"U0_{n}"] = U0[n]
            arrays[f"R_{n}"] = R[n]
            arrays[f"U_{n}"] = U[n]
        np.savez_compressed(HERE / f"{name}.npz", **arrays)
        print(name, {k: v.shape for k, v in arrays.items()})


if __name__ == "__main__":
    main()
