"""Euler2D: the reference's solver API on top of the gfx950 backend.

``Euler2D(config, mesh_config)`` accepts the same ``SolverConfig`` and mesh dictionary / generator
as the reference (solvers/Euler2D.py:46-100, solvers/base.py:44-100).  ``solve()`` follows the
loop of ``Euler2D._solve`` (solvers/Euler2D.py:175-216): dt from the CFL condition clipped to the
final time, one explicit Runge-Kutta step of all blocks, a realizability check, ``t += dt`` --
with every step executed on the device.  ``run_steps(n)`` advances exactly ``n`` steps.
``Euler2DBatch`` advances many independent cases (each with its own dt) in the same launches.
"""
from __future__ import annotations

import logging
import pathlib
from typing import Dict, List, Optional, Sequence

import numpy as np

from .backend import CaseSpec, DevicePlan
from .mesh import MeshGenerator
from .states import ConservativeState, PrimitiveState
from .time_marching import butcher_table

_log = logging.getLogger("pyhype_amd")

SUPPORTED_FVM = ("MUSCL", "MUSCL-HIP")


class _Mesh:
    """The parts of QuadMesh an initial condition reads: centroids shaped (ny, nx, 1)."""

    def __init__(self, geom):
        self.x = geom.xc[:, :, None]
        self.y = geom.yc[:, :, None]
        self.nx, self.ny = geom.nx, geom.ny
        self.A = geom.A[:, :, None]
        self.geometry = geom


class HostBlock:
    """Host-side view of one block for the initial-condition protocol
    (initial_conditions/base.py:28-31): ``mesh.x/y``, ``config``, ``state``."""

    def __init__(self, config, entry):
        self.config = config
        self.global_block_num = entry.num
        self.mesh = _Mesh(entry.geom)
        self.state = ConservativeState(config.fluid, shape=(entry.ny, entry.nx, 4))
        self.entry = entry


def validate_config(config) -> List[List[float]]:
    """The construction-time checks of the reference factories, for the configuration space the
    backend implements (fvm/factory.py:33-65, fvm/SecondOrderMUSCL.py:49-52, flux/factory.py:30-47,
    limiters/factory.py:28-39, gradients/factory.py:26-37, time_marching/factory.py:29-62)."""
    if config.fvm_type not in SUPPORTED_FVM:
        raise ValueError(f"Specified finite volume method type {config.fvm_type} has not been specialized.")
    if config.fvm_spatial_order != 2:
        raise ValueError(f"No MUSCL finite volume method has been specialized with order {config.fvm_spatial_order}")
    if config.nghost != 1:
        raise ValueError("Number of ghost cells must be equal to 1 for this method.")
    if config.fvm_num_quadrature_points not in (1, 2, 3):
        # the _QUADS lookup of QuadraturePointData (mesh/quadratures.py:33-37,71)
        raise KeyError(config.fvm_num_quadrature_points)
    if config.fvm_gradient_type != "GreenGauss":
        raise ValueError(f"Gradient type {config.fvm_gradient_type} is not available.")
    if config.fvm_flux_function_type not in ("Roe", "HLLL"):
        raise ValueError(f"Flux function type {config.fvm_flux_function_type} is not available.")
    if config.fvm_slope_limiter_type not in ("VanLeer", "VanAlbada", "Venkatakrishnan", "BarthJespersen"):
        raise ValueError("MUSCL: Slope limiter type not specified.")
    if config.reconstruction_type not in (ConservativeState, PrimitiveState):
        raise ValueError("reconstruction_type must be ConservativeState or PrimitiveState")
    if config.interface_interpolation != "arithmetic_average":
        raise ValueError("Interface Interpolation method is not defined.")
    if not config.use_JIT:
        # the reference's use_JIT=False path evaluates other formulas (flux/HLLL.py:43-60,
        # flux/base.py:113-117, limiters/limiters.py:33-36); only the JIT arithmetic is implemented
        raise ValueError("use_JIT=False selects the reference's NumPy formulas, which this backend does not "
                         "implement; use use_JIT=True (the reference default)")
    return butcher_table(config.time_integrator, config)


def default_rows(cells: int) -> int:
    """Rows per work item when no tuned value is cached: about 4096 items of 60 columns."""
    return int(max(2, min(64, round(cells / (60 * 4096)))))


def mesh_dict(mesh_config) -> Dict[int, dict]:
    return mesh_config.dict if isinstance(mesh_config, MeshGenerator) else mesh_config


class _HostEntry:
    """The parts of a plan's block entry that HostBlock reads, for host-only set-ups."""

    def __init__(self, num, geom):
        self.num, self.geom, self.nx, self.ny = num, geom, geom.nx, geom.ny


def host_initial_states(config, mesh_config, initial_condition=None) -> Dict[int, np.ndarray]:
    """Initial conservative state of every block, computed on the host only (no device plan): the
    same IC protocol as Euler2D.apply_initial_condition (solvers/Euler2D.py:90-92) on blocks whose
    geometry follows the reference's mesh formulas."""
    from .mesh import BlockGeometry

    ic = config.initial_condition if initial_condition is None else initial_condition
    out = {}
    for n, d in mesh_dict(mesh_config).items():
        geom = BlockGeometry({k: d[k] for k in ("SW", "SE", "NW", "NE")}, config.nx, config.ny)
        hb = HostBlock(config, _HostEntry(n, geom))
        ic.apply_to_block(hb)
        out[n] = np.array(hb.state.data, dtype=np.float64)
    return out


class Euler2DBatch:
    """Independent cases advanced together.  All cases share flux, limiter, time integrator, CFL and
    fluid; each keeps its own block size, dt (CFL of its own blocks) and time."""

    def __init__(self, configs: Sequence, meshes: Sequence, *, rows_per_item: int = 0, grid_waves: int = 0,
                 force_general: bool = False, use_tuning_cache: bool = True):
        if len(configs) != len(meshes) or not configs:
            raise ValueError("need one mesh per config")
        c0 = configs[0]
        self.table = validate_config(c0)
        for c in configs[1:]:
            validate_config(c)
            for attr in ("fvm_flux_function_type", "fvm_slope_limiter_type", "reconstruction_type", "time_integrator",
                         "CFL", "fvm_num_quadrature_points"):
                if getattr(c, attr) != getattr(c0, attr):
                    raise ValueError(f"batched cases must share {attr}")
            if type(c.fluid) is not type(c0.fluid):
                raise ValueError("batched cases must share the fluid")
        self.configs = list(configs)
        self.fluid = c0.fluid
        specs = [CaseSpec(mesh_dict(m), c.nx, c.ny, c.t_final * c.fluid.far_field.a) for c, m in zip(configs, meshes)]
        self.plan = DevicePlan(specs, flux=c0.fvm_flux_function_type, table=self.table, fluid=c0.fluid,
                               cfl=c0.CFL, rows_per_item=rows_per_item, grid_waves=grid_waves,
                               force_general=force_general, limiter=c0.fvm_slope_limiter_type,
                               primitive=c0.reconstruction_type is PrimitiveState,
                               quadrature_points=c0.fvm_num_quadrature_points)
        self.tuned = False
        if rows_per_item == 0:
            from . import autotune

            self.tuned = use_tuning_cache and autotune.apply_cached(self.plan)
            if not self.tuned:
                self.plan.set_launch(default_rows(self.plan.cells()), grid_waves)
        self.case_blocks: List[Dict[int, HostBlock]] = []
        for c, entries in enumerate(self.plan.case_blocks):
            self.case_blocks.append({n: HostBlock(self.configs[c], e) for n, e in entries.items()})
        self.t = np.zeros(len(configs))
        self.num_time_step = 0
        self.failed: Dict[int, int] = {}

    # ---------------------------------------------------------------- set-up
    def apply_initial_condition(self, initial_condition=None):
        """Euler2D.apply_initial_condition (solvers/Euler2D.py:90-92), then upload.  An explicit
        ``initial_condition`` replaces the configured one of every case."""
        for c, blocks in enumerate(self.case_blocks):
            ic = self.configs[c].initial_condition if initial_condition is None else initial_condition
            for blk in blocks.values():
                ic.apply_to_block(blk)
        self.upload()

    def upload(self):
        self.plan.set_states([{n: b.state.data for n, b in blocks.items()} for blocks in self.case_blocks])
        self.plan.reset(self.t, self.num_time_step)

    def set_state(self, states: Sequence[Dict[int, np.ndarray]]):
        for c, blocks in enumerate(self.case_blocks):
            for n, b in blocks.items():
                b.state.data = np.array(states[c][n], dtype=np.float64)
        self.upload()

    # ---------------------------------------------------------------- stepping
    strict = False  # raise on any unrealizable case (True for the single-case Euler2D)

    def run_steps(self, n: int, use_graph: bool = True):
        """Advance exactly n steps (dt per case from the CFL condition, clipped to t_final).

        A case whose state becomes unrealizable (rho <= 0 or E <= 0) is frozen at that step, as
        the reference aborts that run (solvers/Euler2D.py:144-152); ``failed`` maps case -> step.
        With ``strict`` a RealizabilityError is raised instead."""
        if n <= 0:
            return
        try:
            self.plan.run(n, use_graph=use_graph, check=self.strict)
        finally:
            # also after a RealizabilityError: time, step counter and the failing step of each case
            # describe the state the plan holds
            bad, step, t, dt = self.plan.status()
            self.t = t
            self.dt_last = dt
            self.num_time_step = step
            self.failed = {c: int(b) for c, b in enumerate(bad) if b >= 0}

    def autotune(self, steps: int = 10, **kw):
        """Tune the launch shape on the current state (state, time and step are restored)."""
        from . import autotune

        res = autotune.autotune(self.plan, steps=steps, **kw)
        self.tuned = True
        return res

    def states(self) -> List[Dict[int, np.ndarray]]:
        return self.plan.get_states()

    def sync_to_host(self):
        for c, st in enumerate(self.states()):
            for n, arr in st.items():
                self.case_blocks[c][n].state.data = arr

    def last_dt(self) -> np.ndarray:
        return self.plan.status()[3]


class Euler2D(Euler2DBatch):
    """Single-case solver with the reference's constructor, ``solve()`` and file output."""

    strict = True

    def __init__(self, config, mesh_config, **kw):
        super().__init__([config], [mesh_config], **kw)
        self.config = config
        self.t_final = config.t_final * config.fluid.far_field.a
        # Solver.__init__ (solvers/base.py:78-86)
        self.write_path = None
        if getattr(config, "write_solution", False):
            self.write_path = pathlib.Path(config.write_solution_base) / config.write_solution_name
            self.write_path.mkdir(exist_ok=True, parents=True)
            _log.info(self.write_path)

    @property
    def blocks(self):
        return self.case_blocks[0].values()

    def state(self, block_num: int) -> np.ndarray:
        return self.states()[0][block_num]

    # ---------------------------------------------------------------- file output
    def write_mesh(self):
        """Solver.write_mesh (solvers/base.py:142-155): cell centroids of every block as
        ``<write_path>/mesh/mesh_{x,y}_blk_<n>.npy`` with shape (ny, nx, 1)."""
        path = self.write_path / "mesh"
        path.mkdir(parents=True, exist_ok=True)
        for b in self.blocks:
            np.save(str(path / "mesh_x_blk_") + str(b.global_block_num), b.mesh.x)
            np.save(str(path / "mesh_y_blk_") + str(b.global_block_num), b.mesh.y)

    def write_solution(self):
        """Solver.write_solution (solvers/base.py:157-172): in mode ``every_n_timesteps`` and when
        ``num_time_step % write_every_n_timesteps == 0``, the conservative state of every block as
        ``<write_path>/<num_time_step>/<write_solution_name>_blk_<n>.npy`` (ny, nx, 4)."""
        cfg = self.config
        if cfg.write_solution_mode == "every_n_timesteps" and self.num_time_step % cfg.write_every_n_timesteps == 0:
            path = self.write_path / str(self.num_time_step)
            path.mkdir(parents=True, exist_ok=True)
            states = self.plan.get_states()[0]
            for b in self.blocks:
                np.save(str(path / cfg.write_solution_name) + "_blk_" + str(b.global_block_num),
                        states[b.global_block_num])

    def solve(self, chunk: int = 50, max_steps: Optional[int] = None):
        """Run to t_final (solvers/Euler2D.py:97-100,175-216).

        With ``config.write_solution`` the mesh is written first and the solution at the reference's
        points: ``write_solution`` runs after a step's update and before ``num_time_step`` is
        incremented, so the files in directory ``k`` (k a multiple of write_every_n_timesteps) hold
        the state after step k + 1.  Steps between output points run on the device in chunks."""
        self.apply_initial_condition()
        writing = self.write_path is not None
        if writing:
            self.write_mesh()
        every = max(1, int(getattr(self.config, "write_every_n_timesteps", 1) or 1))
        steps = 0
        while self.t[0] < self.t_final and (max_steps is None or steps < max_steps):
            if self.num_time_step % 50 == 0:
                _log.info("Simulation time: %s, Timestep number: %s", self.t[0] / self.fluid.far_field.a,
                          self.num_time_step)
            n = chunk if max_steps is None else min(chunk, max_steps - steps)
            last = float(getattr(self, "dt_last", [0.0])[0])
            if last <= 0.0 or self.t_final - self.t[0] < 1.5 * n * last:
                n = 1  # first step, or near t_final: single steps, so the count stops where the reference's does
            if writing and self.config.write_solution_mode == "every_n_timesteps":
                # the next output point: the step whose pre-increment counter is a multiple of `every`
                k = self.num_time_step
                nxt = k if k % every == 0 else k + (every - k % every)
                n = min(n, nxt - k + 1)
            self.run_steps(n)
            steps += n
            if writing:
                # run_steps advanced num_time_step by n; the reference writes with the counter of the
                # step just taken, before its increment
                self.num_time_step -= 1
                self.write_solution()
                self.num_time_step += 1
        self.sync_to_host()
