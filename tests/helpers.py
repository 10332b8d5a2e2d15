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
                       t_final=config.t_final, a_inf=config.fluid.far_field.a)
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
    return all(np.array_equal(gs[n], rs[n]) for n in rs)
