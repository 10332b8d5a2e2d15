"""The benchmark workloads as SolverConfig + mesh definitions.

* ``explosion_single``: 1 block, 100x100 on [0,10]^2, explosion, Roe/RK4, CFL 0.8, 50 steps.
* ``explosion_multiblock``: 2x4 blocks of 150x150 on [0,10]x[0,20] (RectagularMeshGenerator),
  same physics, 200 steps.
* ``oblique_shock``: 5x1 blocks of 500x500 on [0,5]x[0,1], Air(a_inf=1, rho_inf=1),
  post-shock (8, 8.25cos30, -8.25sin30, 116.5) behind x = 1/6 + y/sqrt(3), pre-shock
  (1.4, 0, 0, 1); Reflection at the bottom, OutletDirichlet elsewhere; HLLL, SSP-RK2, CFL 0.4,
  500 steps.
* ``explosion_batch``: 64 independent single-block cases, side 64 + 32*k with k ~ U{0..14},
  square of width ~U[1,4] at a seeded centre with pressure ratio ~logU[2,10] (density uniform,
  1.1742); Roe/RK4, CFL 0.8, 100 steps, per-case dt.  Scaling the density by the same ratio
  (equal temperature) makes most drawn cases unrealizable within a few steps in the reference
  scheme at CFL 0.8; ``density_scaled=True`` selects that variant.
* ``explosion_adjoint``: 1 block 256x256 on [0,10]^2, explosion, Roe, SSP-RK2, CFL 0.4, 100 steps;
  ``adjoint_problem`` adds the differentiable base state and the density-bump target of the
  discrete-adjoint objective J = sum (rho(T) - rho_target(T))^2.

Every function accepts a scale-down (``n``, ``nb``...) so tests can run the same set-up small.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

from .fluids import Air
from .initial_conditions import (DensityBumpIC, DoubleMachReflectionIC, ExplosionInitialCondition, ImplosionIC,
                                 ObliqueShockIC, ShockBoxIC, SmoothExplosionIC, SquareExplosionIC)
from .mesh import QuadMeshGenerator, RectagularMeshGenerator
from .solver_config import SolverConfig
from .states import PrimitiveState

T_FINAL_NONE = 1e30   # step-count runs: t_final never clips dt


def single_block_mesh(east=10.0, north=10.0, bc="Reflection"):
    return {0: {
        "nBLK": 0, "NW": [0, north], "NE": [east, north], "SW": [0, 0], "SE": [east, 0],
        "NeighborE": None, "NeighborW": None, "NeighborN": None, "NeighborS": None,
        "NeighborNE": None, "NeighborNW": None, "NeighborSE": None, "NeighborSW": None,
        "BCTypeE": bc, "BCTypeW": bc, "BCTypeN": bc, "BCTypeS": bc,
        "BCTypeNE": None, "BCTypeNW": None, "BCTypeSE": None, "BCTypeSW": None,
    }}


def _config(n, m, *, ic, flux="Roe", integrator="RK4", cfl=0.8, fluid=None, alpha=None):
    return SolverConfig(
        nx=n, ny=m, CFL=cfl, t_final=T_FINAL_NONE, initial_condition=ic, fvm_type="MUSCL-HIP",
        time_integrator=integrator, fvm_gradient_type="GreenGauss", fvm_flux_function_type=flux,
        fvm_slope_limiter_type="Venkatakrishnan", fvm_spatial_order=2, fvm_num_quadrature_points=1,
        fluid=fluid if fluid is not None else Air(a_inf=343.0, rho_inf=1.0), nghost=1, alpha=alpha)


def explosion_single(n: int = 100):
    return _config(n, n, ic=ExplosionInitialCondition()), single_block_mesh()


def explosion_multiblock(n: int = 150, nbx: int = 2, nby: int = 4):
    mesh = RectagularMeshGenerator.generate(BCE=["Reflection"], BCW=["Reflection"], BCN=["Reflection"],
                                            BCS=["Reflection"], east=10.0, west=0.0, north=20.0, south=0.0,
                                            n_blocks_horizontal=nbx, n_blocks_vertical=nby)
    return _config(n, n, ic=ExplosionInitialCondition()), mesh


def oblique_shock(n: int = 500, nblocks: int = 5):
    mesh = RectagularMeshGenerator.generate(BCE=["OutletDirichlet"], BCW=["OutletDirichlet"],
                                            BCN=["OutletDirichlet"], BCS=["Reflection"], east=float(nblocks),
                                            west=0.0, north=1.0, south=0.0, n_blocks_horizontal=nblocks,
                                            n_blocks_vertical=1)
    cfg = _config(n, n, ic=ObliqueShockIC(), flux="HLLL", integrator="SSPRK2", cfl=0.4,
                  fluid=Air(a_inf=1.0, rho_inf=1.0))
    return cfg, mesh


def double_mach_reflection(n: int = 50):
    """The reference's double-Mach-reflection example: 4 x 1 blocks of n x n cells over a 30 degree
    ramp (examples/dmr/mesh.py:1-33: transfinite QuadMeshGenerator mesh, outlets, slip walls on the
    ramp), HLLL, RK2, PrimitiveState reconstruction, Venkatakrishnan, CFL 0.4
    (examples/dmr/config.py:8-34)."""
    k = 1
    a = 2 / np.sqrt(3)
    d = np.tan(30 * np.pi / 180)
    x = [0, k, 2 * k, 3 * k, 4 * k]
    gen = QuadMeshGenerator(nx_blk=4, ny_blk=1, BCE=["OutletDirichlet"], BCW=["OutletDirichlet"],
                            BCN=["OutletDirichlet"], BCS=["OutletDirichlet", "Slipwall", "Slipwall", "Slipwall"],
                            top_x=x, bot_x=x, top_y=[a, a, a + d, a + 2 * d, a + 3 * d],
                            bot_y=[0, 0, d, 2 * d, 3 * d], left_x=[0, 0], right_x=[4 * k, 4 * k],
                            left_y=[0, a], right_y=[3 * d, a + 3 * d])
    cfg = _config(n, n, ic=DoubleMachReflectionIC(), flux="HLLL", integrator="RK2", cfl=0.4)
    cfg.reconstruction_type = PrimitiveState
    return cfg, gen.dict


def implosion(n: int = 40):
    """The reference's implosion example: one 10 x 10 block with reflecting walls, Roe, RK2, CFL 0.4
    (examples/implosion/config.py, mesh.py)."""
    return _config(n, n, ic=ImplosionIC(), integrator="RK2", cfl=0.4), single_block_mesh()


def shockbox(n: int = 50):
    """The reference's shock-box example: one 10 x 10 block with reflecting walls, Roe, RK2, CFL 0.4
    (examples/shockbox/config.py, mesh.py)."""
    return _config(n, n, ic=ShockBoxIC(), integrator="RK2", cfl=0.4), single_block_mesh()


def explosion_batch(seed: int = 0, ncases: int = 64, sizes=None,
                    density_scaled: bool = False) -> Tuple[List[SolverConfig], List[dict], dict]:
    """Configs and meshes of the batched uneven cases, plus the drawn parameters.  Case c's draws
    (side k, width, centre, ratio) come from one seeded stream in case order, so the first cases
    of a smaller batch equal those of a larger one; ``sizes`` overrides the drawn sides."""
    rng = np.random.default_rng(seed)
    configs, meshes, params = [], [], []
    for c in range(ncases):
        k = int(rng.integers(0, 15))
        n = 64 + 32 * k if sizes is None else int(sizes[c])
        w = float(rng.uniform(1.0, 4.0))
        cx, cy = (float(v) for v in rng.uniform(w / 2, 10 - w / 2, size=2))
        r = float(np.exp(rng.uniform(np.log(2.0), np.log(10.0))))
        ic = SquareExplosionIC(cx - w / 2, cx + w / 2, cy - w / 2, cy + w / 2, r,
                               density_ratio=r if density_scaled else 1.0)
        configs.append(_config(n, n, ic=ic))
        meshes.append(single_block_mesh())
        params.append({"n": n, "w": w, "cx": cx, "cy": cy, "ratio": r})
    return configs, meshes, {"seed": seed, "cases": params}


def explosion_adjoint(n: int = 256):
    return _config(n, n, ic=ExplosionInitialCondition(), integrator="SSPRK2", cfl=0.4), single_block_mesh()


def adjoint_problem(n: int = 256, seed: int = 0, width: float = 0.1, amplitude: float = 0.01):
    """Set-up of the discrete-adjoint workload on ``explosion_adjoint``'s mesh and scheme: the
    linearisation point is the smoothed explosion (``SmoothExplosionIC``, background phases drawn
    from ``seed``); the target is the same state with a Gaussian density bump of relative
    amplitude 1e-2 and sigma 1 at a centre drawn from ``seed``.  Returns (config, mesh, target IC,
    parameters)."""
    rng = np.random.default_rng(seed)
    phases = tuple(float(p) for p in rng.uniform(0.0, 2.0 * np.pi, 6))
    cx, cy = (float(v) for v in rng.uniform(2.0, 8.0, 2))
    base = SmoothExplosionIC(width=width, amplitude=amplitude, phases=phases)
    target = DensityBumpIC(base, 1e-2, cx, cy, 1.0)
    cfg, mesh = explosion_adjoint(n)
    cfg.initial_condition = base
    return cfg, mesh, target, {"seed": seed, "phases": phases, "bump_center": (cx, cy), "bump_sigma": 1.0,
                               "bump_amplitude": 1e-2, "width": width, "amplitude": amplitude}


WORKLOADS = {
    "explosion_single": (explosion_single, 50, 4),
    "explosion_multiblock": (explosion_multiblock, 200, 4),
    "oblique_shock": (oblique_shock, 500, 2),
    "explosion_batch": (explosion_batch, 100, 4),
    "explosion_adjoint": (explosion_adjoint, 100, 2),
    "double_mach_reflection": (double_mach_reflection, 100, 2),
}
