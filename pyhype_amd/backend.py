"""Device plan: block table, packed geometry and the C-ABI calls, with torch as the buffer carrier.

``DevicePlan`` turns one or more independent cases (each a reference-style block dictionary with
its own ``nx``/``ny``) into a ``phx_plan`` of the HIP library and owns the FP64 state tensor on the
GPU.  All arithmetic runs in the library's kernels; this module only lays out memory and calls
them.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _native as nat
from .mesh import BlockGeometry
from .states import RealizabilityError

_WALLS = ("Reflection", "Slipwall")


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("pyhype_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


@dataclass
class CaseSpec:
    """One independent simulation: block dictionaries (mesh/base.py:201-223) and block size."""
    mesh: Dict[int, dict]
    nx: int
    ny: int
    t_final: float = float("inf")   # already scaled by a_inf


@dataclass
class BlockEntry:
    case: int
    num: int            # block number within its case (nBLK)
    index: int          # global block index in the plan
    geom: BlockGeometry
    offset: int         # state offset in doubles
    nx: int
    ny: int
    sides: List[int]
    neighbors: List[int]
    cartesian: bool


def stage_programs(table) -> List[dict]:
    """The accumulator-form stage programs the library derives from a Butcher table (same rule as
    build_programs in csrc/phx_kernels.hip): buffer names "U" (state / U0), "S0", "S1" (stage
    states) and "A<m>" (accumulators)."""
    S = len(table)
    P = ["U"] * S
    free, nacc, progs = [], 0, []
    for s in range(S):
        inp = "U" if s == 0 else ("S0" if (s - 1) % 2 == 0 else "S1")
        out = "U" if (s == S - 1 and S > 1) else ("S0" if s % 2 == 0 else "S1")
        outs = [{"dst": out, "src": P[s], "a": float(table[s][s])}]
        if P[s].startswith("A"):
            free.append(P[s])
        for t in range(s + 1, S):
            c = float(table[t][s])
            if c == 0.0:
                continue
            if P[t] == "U":
                name = free.pop() if free else f"A{nacc}"
                if name == f"A{nacc}":
                    nacc += 1
                outs.append({"dst": name, "src": "U", "a": c})
                P[t] = name
            else:
                outs.append({"dst": P[t], "src": P[t], "a": c})
        progs.append({"in": inp, "outs": outs})
    return progs


def side_kind(bc, neighbor) -> int:
    """Ghost/outer-state rule of a side (blocks/ghost.py:128-143,187-216; fvm/base.py:305-325)."""
    if bc is None:
        return nat.SIDE_OPEN if neighbor is None else nat.SIDE_NEIGHBOR
    if bc in _WALLS:
        return nat.SIDE_WALL
    if bc == "OutletDirichlet":
        return nat.SIDE_OUTLET
    raise ValueError("Boundary Condition type " + str(bc) + " has not been specialized.")


class Layout:
    """Blocks of every case in state order, with their packed geometry and C descriptors."""

    def __init__(self, blocks, case_blocks, state_len, geoms, descs):
        self.blocks, self.case_blocks, self.state_len = blocks, case_blocks, state_len
        self.geoms, self.descs = geoms, descs


def layout_cases(cases: Sequence[CaseSpec], force_general: bool = False, quadrature_points: int = 1) -> Layout:
    """State layout and ``phx_block_desc`` array of a list of cases (host only, no GPU needed).
    Blocks of a case are numbered in mesh-dictionary order; each block's state is its reference
    ``(ny, nx, 4)`` array, blocks concatenated.  With ``quadrature_points`` > 1 each block's packed
    geometry ends with its Gauss-point offsets."""
    blocks: List[BlockEntry] = []
    case_blocks: List[Dict[int, BlockEntry]] = []
    offset = 0
    for c, spec in enumerate(cases):
        index_of = {}
        entries = {}
        for n, d in spec.mesh.items():
            g = BlockGeometry({k: d[k] for k in ("SW", "SE", "NW", "NE")}, spec.nx, spec.ny)
            e = BlockEntry(c, n, len(blocks), g, offset, spec.nx, spec.ny, [], [], False)
            offset += 4 * spec.nx * spec.ny
            index_of[n] = e.index
            blocks.append(e)
            entries[n] = e
        for n, d in spec.mesh.items():
            e = entries[n]
            for s in ("E", "W", "N", "S"):
                bc, nb = d["BCType" + s], d["Neighbor" + s]
                k = side_kind(bc, nb)
                if k == nat.SIDE_NEIGHBOR and nb not in index_of:
                    raise ValueError(f"block {n}: neighbour {nb} does not exist")
                e.sides.append(k)
                e.neighbors.append(index_of[nb] if k == nat.SIDE_NEIGHBOR else -1)
        case_blocks.append(entries)
    geoms = []
    descs = (nat.BlockDesc * len(blocks))()
    for e, dsc in zip(blocks, descs):
        packed = None if force_general else e.geom.pack_cartesian()
        e.cartesian = packed is not None
        if packed is None:
            packed = e.geom.pack_general()
        if quadrature_points > 1:
            packed = np.concatenate([packed, e.geom.pack_quadrature(quadrature_points)])
        packed = np.ascontiguousarray(packed, dtype=np.float64)
        geoms.append(packed)
        dsc.nx, dsc.ny, dsc.case_id = e.nx, e.ny, e.case
        dsc.geom_kind = nat.GEOM_CARTESIAN if e.cartesian else nat.GEOM_GENERAL
        for s in range(4):
            dsc.side_type[s] = e.sides[s]
            dsc.neighbor[s] = e.neighbors[s]
        dsc.state_offset = e.offset
        dsc.geom = packed.ctypes.data_as(C.POINTER(C.c_double))
        dsc.geom_len = packed.size
    return Layout(blocks, case_blocks, offset, geoms, descs)


class DevicePlan:
    def __init__(self, cases: Sequence[CaseSpec], *, flux: str, table: List[List[float]], fluid, cfl: float,
                 rows_per_item: int = 0, grid_waves: int = 0, force_general: bool = False,
                 limiter: str = "Venkatakrishnan", primitive: bool = False, quadrature_points: int = 1):
        torch = _torch()
        self.torch = torch
        self.lib = nat.load()
        self.cases = list(cases)
        flux_id = {"Roe": nat.FLUX_ROE, "HLLL": nat.FLUX_HLLL}.get(flux)
        if flux_id is None:
            raise ValueError(f"Flux function type {flux} is not available.")
        self.flux = flux
        if limiter not in nat.LIMITERS:
            raise ValueError("MUSCL: Slope limiter type not specified.")
        self.limiter = limiter
        self.primitive = bool(primitive)
        if quadrature_points not in (1, 2, 3):
            raise KeyError(quadrature_points)   # _QUADS[...] lookup (mesh/quadratures.py:71)
        self.quadrature_points = int(quadrature_points)
        self.table = table
        S = len(table)
        a = np.zeros((S, S))
        for s, row in enumerate(table):
            if len(row) != s + 1:
                raise ValueError("Butcher table rows must have stage+1 entries")
            a[s, : s + 1] = row
        self._a = np.ascontiguousarray(a)
        self.nstages = S
        lay = layout_cases(self.cases, force_general=force_general, quadrature_points=self.quadrature_points)
        self.blocks, self.case_blocks, self.state_len = lay.blocks, lay.case_blocks, lay.state_len
        self._geoms, descs = lay.geoms, lay.descs
        self._tfinal = np.array([c.t_final for c in self.cases], dtype=np.float64)
        g = fluid.gamma()
        pd = nat.PlanDesc()
        pd.abi_version = nat.PHX_ABI_VERSION
        pd.nblocks = len(self.blocks)
        pd.blocks = descs
        pd.ncases = len(self.cases)
        pd.flux = flux_id
        pd.limiter = nat.LIMITERS[limiter]
        pd.reconstruction = nat.RECON_PRIMITIVE if primitive else nat.RECON_CONSERVATIVE
        pd.quadrature_points = self.quadrature_points
        pd.nstages = S
        pd.butcher_a = self._a.ctypes.data_as(C.POINTER(C.c_double))
        pd.gamma = g
        pd.gm1 = fluid.gm1()
        pd.one_over_gm1 = fluid.one_over_gm1()
        pd.g_over_gm1 = fluid.g_over_gm1()
        pd.cfl = cfl
        pd.t_final = self._tfinal.ctypes.data_as(C.POINTER(C.c_double))
        pd.rows_per_item = rows_per_item
        pd.grid_waves = grid_waves
        self._descs = descs
        handle = C.c_void_p()
        nat.check(self.lib.phx_plan_create(C.byref(pd), C.byref(handle)), "phx_plan_create")
        self.handle = handle
        self.U = torch.zeros(self.state_len, dtype=torch.float64, device="cuda")
        self.nsteps_done = 0
        self.launch = {"rows_per_item": rows_per_item or 32, "grid_waves": grid_waves, "waves_per_simd": 2}
        if self.quadrature_points > 1:   # the plan runs a split stage (one y-face per wave, 3 rows per item)
            self.launch.update({"variant": 6, "rows_b": 3})

    # ------------------------------------------------------------------ lifetime
    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.phx_plan_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ helpers
    def stream(self):
        return C.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    @staticmethod
    def ptr(t) -> C.c_void_p:
        return C.c_void_p(t.data_ptr())

    def block_view(self, e: BlockEntry, buf=None):
        buf = self.U if buf is None else buf
        return buf[e.offset:e.offset + 4 * e.nx * e.ny].view(e.ny, e.nx, 4)

    def set_states(self, states: Dict[int, Dict[int, np.ndarray]]):
        """states[case][block_num] = (ny, nx, 4) conservative array."""
        host = np.zeros(self.state_len)
        for e in self.blocks:
            arr = np.asarray(states[e.case][e.num], dtype=np.float64)
            if arr.shape != (e.ny, e.nx, 4):
                raise ValueError(f"state of block {e.num} has shape {arr.shape}, expected {(e.ny, e.nx, 4)}")
            host[e.offset:e.offset + arr.size] = arr.ravel()
        self.U.copy_(self.torch.from_numpy(host))

    def get_states(self, buf=None) -> List[Dict[int, np.ndarray]]:
        host = (self.U if buf is None else buf).cpu().numpy()
        out: List[Dict[int, np.ndarray]] = [dict() for _ in self.cases]
        for e in self.blocks:
            out[e.case][e.num] = host[e.offset:e.offset + 4 * e.nx * e.ny].reshape(e.ny, e.nx, 4).copy()
        return out

    # ------------------------------------------------------------------ operations
    def reset(self, t0: Optional[Sequence[float]] = None, step0: int = 0):
        t = np.zeros(len(self.cases)) if t0 is None else np.asarray(t0, dtype=np.float64)
        nat.check(self.lib.phx_reset(self.handle, t.ctypes.data_as(C.POINTER(C.c_double)), step0, self.stream()),
                  "phx_reset")

    def run(self, nsteps: int, use_graph: bool = False, check: bool = True):
        """Advance nsteps; with check, raise RealizabilityError if any case failed."""
        nat.check(self.lib.phx_run(self.handle, self.ptr(self.U), int(nsteps), int(bool(use_graph)), self.stream()),
                  "phx_run")
        self.nsteps_done += nsteps
        if check:
            self.check_realizable()

    def run_frozen(self, dt_sched, check: bool = True):
        """dt_sched: device tensor (nsteps, ncases)."""
        nsteps = int(dt_sched.shape[0])
        nat.check(self.lib.phx_run_frozen(self.handle, self.ptr(self.U), nsteps, self.ptr(dt_sched), self.stream()),
                  "phx_run_frozen")
        self.nsteps_done += nsteps
        if check:
            self.check_realizable()

    def stage(self, s: int):
        """One Runge-Kutta stage of the plan's program on the current state (for timing)."""
        nat.check(self.lib.phx_stage(self.handle, self.ptr(self.U), int(s), self.stream()), "phx_stage")

    def status(self):
        """(bad_step per case (-1 = fine), steps taken, time per case, last dt per case)."""
        nc = len(self.cases)
        bad = (C.c_int32 * nc)()
        step = C.c_int32()
        t = np.zeros(nc)
        dt = np.zeros(nc)
        nat.check(self.lib.phx_status(self.handle, bad, C.byref(step),
                                      t.ctypes.data_as(C.POINTER(C.c_double)),
                                      dt.ctypes.data_as(C.POINTER(C.c_double))), "phx_status")
        return np.array(bad[:], dtype=np.int64), step.value, t, dt

    def failed_cases(self):
        bad = self.status()[0]
        return {c: int(b) for c, b in enumerate(bad) if b >= 0}

    def check_realizable(self):
        failed = self.failed_cases()
        if failed:
            c, b = min(failed.items(), key=lambda kv: kv[1])
            raise RealizabilityError(
                f"ConservativeState has unrealizable values (rho <= 0 or E <= 0) in case {c} after step {b}"
                + (f"; failed cases: {failed}" if len(failed) > 1 else ""))

    def residual(self, U=None):
        torch = self.torch
        U = self.U if U is None else U
        R = torch.zeros_like(U)
        nat.check(self.lib.phx_residual(self.handle, self.ptr(U), self.ptr(R), self.stream()), "phx_residual")
        return R

    def dt(self, U=None):
        U = self.U if U is None else U
        out = self.torch.zeros(len(self.cases), dtype=self.torch.float64, device="cuda")
        nat.check(self.lib.phx_dt(self.handle, self.ptr(U), self.ptr(out), self.stream()), "phx_dt")
        return out

    def ghosts(self, U=None):
        """Ghost strips per block: dict (case, num) -> {E: (ny,1,4), W, N: (1,nx,4), S}."""
        U = self.U if U is None else U
        n = sum(4 * (2 * e.ny + 2 * e.nx) for e in self.blocks)
        G = self.torch.zeros(n, dtype=self.torch.float64, device="cuda")
        nat.check(self.lib.phx_ghosts(self.handle, self.ptr(U), self.ptr(G), self.stream()), "phx_ghosts")
        g = G.cpu().numpy()
        out, off = {}, 0
        for e in self.blocks:
            m = 2 * e.ny + 2 * e.nx
            a = g[off:off + 4 * m].reshape(m, 4)
            off += 4 * m
            out[(e.case, e.num)] = {
                "E": a[:e.ny].reshape(e.ny, 1, 4), "W": a[e.ny:2 * e.ny].reshape(e.ny, 1, 4),
                "N": a[2 * e.ny:2 * e.ny + e.nx].reshape(1, e.nx, 4), "S": a[2 * e.ny + e.nx:].reshape(1, e.nx, 4),
            }
        return out

    def residual_vjp(self, U, lamR):
        """(dR/dU)^T lamR at state U (device tensors of the state layout)."""
        lamU = self.torch.zeros_like(U)
        nat.check(self.lib.phx_residual_vjp(self.handle, self.ptr(U), self.ptr(lamR), self.ptr(lamU), self.stream()),
                  "phx_residual_vjp")
        return lamU

    def residual_jvp(self, U, dU):
        """(dR/dU) dU at state U (device tensors of the state layout)."""
        dR = self.torch.zeros_like(U)
        nat.check(self.lib.phx_residual_jvp(self.handle, self.ptr(U), self.ptr(dU), self.ptr(dR), self.stream()),
                  "phx_residual_jvp")
        return dR

    def step_jvp(self, stages, dt, dU):
        """In place: dU (derivative of the state before one step) -> derivative after it.
        stages: (nstages, state_len) stage inputs of that step; dt: (ncases,) its time steps."""
        nat.check(self.lib.phx_step_jvp(self.handle, self.ptr(stages), self.ptr(dt), self.ptr(dU), self.stream()),
                  "phx_step_jvp")

    def run_record(self, dt_sched, stages):
        """Frozen-dt steps on the current state that record every stage input into ``stages``
        (device tensor (nsteps, nstages, state_len))."""
        nsteps = int(dt_sched.shape[0])
        if tuple(stages.shape) != (nsteps, self.nstages, self.state_len):
            raise ValueError(f"stages must have shape {(nsteps, self.nstages, self.state_len)}")
        nat.check(self.lib.phx_run_record(self.handle, self.ptr(self.U), nsteps, self.ptr(dt_sched),
                                          self.ptr(stages), self.stream()), "phx_run_record")
        self.nsteps_done += nsteps

    def step_vjp(self, stages, dt, lam):
        """In place: lam (adjoint of the state after one step) -> adjoint of the state before it.
        stages: (nstages, state_len) stage inputs of that step; dt: (ncases,) its time steps."""
        nat.check(self.lib.phx_step_vjp(self.handle, self.ptr(stages), self.ptr(dt), self.ptr(lam), self.stream()),
                  "phx_step_vjp")

    def set_launch(self, rows_per_item: int, grid_waves: int = 0, waves_per_simd: int = 2,
                   variant: Optional[int] = None, rows_b: int = 0, xcd_chunk: Optional[int] = None):
        """Launch shape of the stage: rows per work item, persistent grid (0 = one wave per item),
        register budget in waves per SIMD; optionally the stage variant (0 fused, 1 split, 2 shared
        halo, 3 element-wise, 4 paired, 5 split with fused y-face/update, 6 the same with one face per
        wave, 7 variant 6 with the row-major reconstruction kernel, 8 that reconstruction with the
        y-faces and update over flat cell ranges), the rows per item of the split variants' update kernel
        and the order of the split variants' work lists over the XCDs (0 contiguous ranges, k > 0
        chunks of k items dealt in turn)."""
        if xcd_chunk is not None:
            nat.check(self.lib.phx_set_xcd_order(self.handle, int(xcd_chunk)), "phx_set_xcd_order")
        if variant is not None:
            nat.check(self.lib.phx_set_variant(self.handle, int(variant), int(rows_b)), "phx_set_variant")
        nat.check(self.lib.phx_set_launch(self.handle, int(rows_per_item), int(grid_waves), int(waves_per_simd)),
                  "phx_set_launch")
        launch = dict(getattr(self, "launch", {}))
        launch.update({"rows_per_item": int(rows_per_item), "grid_waves": int(grid_waves),
                       "waves_per_simd": int(waves_per_simd)})
        if variant is not None:
            launch["variant"] = int(variant)
            if rows_b:
                launch["rows_b"] = int(rows_b)
        if xcd_chunk is not None:
            launch["xcd_chunk"] = int(xcd_chunk)
        self.launch = launch

    def launch_info(self):
        v = [C.c_int32() for _ in range(4)]
        nat.check(self.lib.phx_launch_info(self.handle, *[C.byref(x) for x in v]), "phx_launch_info")
        out = {"items_cart": v[0].value, "items_gen": v[1].value, "grid_cart": v[2].value, "grid_gen": v[3].value}
        out.update(getattr(self, "launch", {}))
        return out

    def cells(self) -> int:
        return sum(e.nx * e.ny for e in self.blocks)
