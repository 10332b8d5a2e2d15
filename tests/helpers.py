"""Shared helpers of the parity tests: run the oracle on the same inputs as the HIP backend."""
from __future__ import annotations

import numpy as np

from oracle import pyhype_oracle as O
from pyhype_amd.solvers import mesh_dict
from pyhype_amd.time_marching import butcher_table


def oracle_for(config, mesh, states0):
    """Oracle Simulation with the same blocks, scheme and initial state as a pyhype_amd solver."""
    sim = O.Simulation(mesh_dict(mesh), config.nx, config.ny, flux=config.fvm_flux_function_type,
                       integrator=butcher_table(config.time_integrator, config), CFL=config.CFL,
                       t_final=config.t_final, a_inf=config.fluid.far_field.a,
                       limiter=config.fvm_slope_limiter_type,
                       reconstruction="Primitive" if config.reconstruction_type.__name__ == "PrimitiveState"
                       else "Conservative", quadrature_points=config.fvm_num_quadrature_points)
    sim.set_state({n: np.array(a) for n, a in states0.items()})
    return sim


def rel_err(g, r):
    """Per-variable max|g - r| / max|r| (the parity metric), maximised over variables."""
    g = np.asarray(g)
    r = np.asarray(r)
    worst = 0.0
    for k in range(4):
        den = np.max(np.abs(r[..., k]))
        num = np.max(np.abs(g[..., k] - r[..., k]))
        worst = max(worst, num / den if den > 0 else num)
    return worst


def max_rel_err_blocks(gs, rs):
    return max(rel_err(gs[n], rs[n]) for n in rs)


def bitwise_equal_blocks(gs, rs):
    """Equal values in every block (NaNs of a case that became unrealizable compare equal when they
    sit at the same positions)."""
    return all(np.array_equal(gs[n], rs[n], equal_nan=True) for n in rs)


def skewed_mesh():
    """Two genuinely non-rectangular blocks sharing a slanted edge (a wedge-like channel): west
    outlet, east wall, walls north and south."""
    def blk(n, sw, se, nw, ne, nbE, nbW, bcE, bcW):
        return {"nBLK": n, "NW": nw, "NE": ne, "SW": sw, "SE": se,
                "NeighborE": nbE, "NeighborW": nbW, "NeighborN": None, "NeighborS": None,
                "NeighborNE": None, "NeighborNW": None, "NeighborSE": None, "NeighborSW": None,
                "BCTypeE": bcE, "BCTypeW": bcW, "BCTypeN": "Reflection", "BCTypeS": "Reflection",
                "BCTypeNE": None, "BCTypeNW": None, "BCTypeSE": None, "BCTypeSW": None}
    return {0: blk(0, [0.0, 0.0], [5.0, 0.8], [0.3, 10.0], [5.6, 9.1], 1, None, None, "OutletDirichlet"),
            1: blk(1, [5.0, 0.8], [10.0, 1.9], [5.6, 9.1], [10.4, 8.0], None, 0, "Reflection", None)}
