"""Discrete adjoint of the explicit Runge-Kutta trajectory, behind ``torch.autograd.Function``.

The reference has no adjoint.  This module differentiates the exact discrete map of the backend:
``n`` Runge-Kutta steps with a frozen time-step schedule (the schedule of a plain forward run, so
that dt does not depend on the state being differentiated), U_0 -> U_n.  The reverse sweep uses
binomial-free uniform checkpointing: the forward pass keeps U every ``checkpoint_every`` steps; the
reverse pass re-runs each segment from its checkpoint while recording the stage inputs
(``phx_run_record``), then applies the step VJP of every step of the segment in reverse order
(``phx_step_vjp``); the re-run of the next segment overlaps the reverse sweep of the current one
on a second stream.  All arithmetic is in the HIP library; torch holds the buffers and chains the
result into user objectives.

    traj = Trajectory(plan, dt_sched, checkpoint_every=10)
    U_T = trajectory(U0, traj)                # differentiable in U0
    J = ((rho(U_T) - rho_target) ** 2).sum(); J.backward()
    dU_T = traj.jvp(U0, dU0)                  # forward mode: (dU_T/dU_0) dU0

The forward-mode product (``Trajectory.jvp``, ``phx_step_jvp``) runs the residual's own arithmetic
on value-and-derivative pairs; <J v, w> = <v, J^T w> ties it to the reverse sweep.
"""
from __future__ import annotations

from typing import Optional

import numpy as np


def frozen_schedule(plan, nsteps: int):
    """Advance ``plan`` by ``nsteps`` CFL steps from its current state and return the dt of every
    step as a device tensor (nsteps, ncases) (one host synchronisation per step)."""
    torch = plan.torch
    dts = np.zeros((nsteps, len(plan.cases)))
    for s in range(nsteps):
        plan.run(1, use_graph=False, check=True)
        dts[s] = plan.status()[3]
    return torch.as_tensor(dts, dtype=torch.float64, device="cuda")


class Trajectory:
    """``len(dt_sched)`` frozen-dt steps of ``plan`` as a map of the state, with checkpoints."""

    RECORD_BUFFERS = 3  # segments whose recorded stage inputs can be held at once in the reverse sweep

    def __init__(self, plan, dt_sched, checkpoint_every: int = 10):
        if checkpoint_every < 1:
            raise ValueError("checkpoint_every must be >= 1")
        self.plan = plan
        self.dt = dt_sched.contiguous()
        self.nsteps = int(self.dt.shape[0])
        self.every = int(checkpoint_every)
        self.checkpoints = []
        torch = plan.torch
        m = min(self.every, max(1, self.nsteps))
        self.stages = torch.empty((m, plan.nstages, plan.state_len), dtype=torch.float64, device="cuda")
        self._stages_more = []  # further record buffers, allocated by the first backward with > 1 segment
        self._side = None       # stream of the segment recomputation (default priority)

    def forward(self, U0):
        """U_n from U_0 (device tensor of the state layout); keeps the checkpoints."""
        P = self.plan
        P.U.copy_(U0)
        self.checkpoints = []
        for s0 in range(0, self.nsteps, self.every):
            self.checkpoints.append(P.U.clone())
            P.run_frozen(self.dt[s0:s0 + self.every], check=False)
        P.check_realizable()
        return P.U.clone()

    def backward(self, lam_T):
        """Adjoint of U_0 from the adjoint of U_n (the checkpoints of the last forward).

        The recomputation of the next segments (``run_record`` from their checkpoints, on a second
        stream) overlaps the reverse sweep of the current one (``step_vjp``, on the caller's
        stream): the two use disjoint device buffers (the plan's state and scratch versus its
        adjoint work arrays), and the records rotate through ``RECORD_BUFFERS`` buffers, each
        reused only after the sweep that read it has finished (events), so the recomputation can
        run up to two segments ahead of the sweep."""
        if len(self.checkpoints) != (self.nsteps + self.every - 1) // self.every:
            raise RuntimeError("Trajectory.backward needs a preceding forward")
        P = self.plan
        torch = P.torch
        lam = lam_T.detach().to(dtype=torch.float64).contiguous().clone()
        segs = list(reversed(range(len(self.checkpoints))))

        def span(seg):
            s0 = seg * self.every
            return s0, min(self.every, self.nsteps - s0)

        if len(segs) == 1:
            s0, m = span(segs[0])
            P.U.copy_(self.checkpoints[segs[0]])
            P.run_record(self.dt[s0:s0 + m], self.stages[:m])
            for i in reversed(range(m)):
                P.step_vjp(self.stages[i], self.dt[s0 + i], lam)
            return lam

        nbuf = min(self.RECORD_BUFFERS, len(segs))
        while len(self._stages_more) < nbuf - 1:
            self._stages_more.append(torch.empty_like(self.stages))
        if self._side is None:
            self._side = torch.cuda.Stream()
        main, side = torch.cuda.current_stream(), self._side
        bufs = [self.stages] + self._stages_more[:nbuf - 1]
        ready = [torch.cuda.Event() for _ in range(nbuf)]
        free = [torch.cuda.Event() for _ in range(nbuf)]

        def recompute(idx):
            """Re-run segment segs[idx] into record buffer idx % nbuf, once the sweep that read that
            buffer before (segment segs[idx - nbuf]) has finished."""
            s0, m = span(segs[idx])
            b = idx % nbuf
            if idx >= nbuf:
                side.wait_event(free[b])
            with torch.cuda.stream(side):
                P.U.copy_(self.checkpoints[segs[idx]])
                P.run_record(self.dt[s0:s0 + m], bufs[b][:m])
                ready[b].record(side)

        side.wait_stream(main)          # checkpoints, schedule and state written on the caller's stream
        for idx in range(nbuf - 1):
            recompute(idx)
        for k, seg in enumerate(segs):
            if k + nbuf - 1 < len(segs):
                recompute(k + nbuf - 1)
            b = k % nbuf
            main.wait_event(ready[b])
            s0, m = span(seg)
            for i in reversed(range(m)):
                P.step_vjp(bufs[b][i], self.dt[s0 + i], lam)
            free[b].record(main)
        main.wait_stream(side)          # the plan's state is the caller's again
        return lam


    def jvp(self, U0, dU0):
        """Forward-mode product of the trajectory: the derivative of U_n in direction dU0 at U0.
        Each segment is run from U0 with its stage inputs recorded (``phx_run_record``), then the
        step JVPs carry dU through the segment's steps (``phx_step_jvp``).  The plan's state ends
        at U_n."""
        P = self.plan
        torch = P.torch
        dU = dU0.detach().to(dtype=torch.float64).contiguous().clone()
        P.U.copy_(U0)
        for s0 in range(0, self.nsteps, self.every):
            m = min(self.every, self.nsteps - s0)
            P.run_record(self.dt[s0:s0 + m], self.stages[:m])
            for i in range(m):
                P.step_jvp(self.stages[i], self.dt[s0 + i], dU)
        return dU


_FN = None


def _function():
    global _FN
    if _FN is None:
        import torch

        class TrajectoryFunction(torch.autograd.Function):
            """U_0 -> U_n of a ``Trajectory``; backward is the discrete adjoint."""

            @staticmethod
            def forward(ctx, U0, traj):
                ctx.traj = traj
                return traj.forward(U0.detach())

            @staticmethod
            def backward(ctx, grad_out):
                return ctx.traj.backward(grad_out), None

        _FN = TrajectoryFunction
    return _FN


def trajectory(U0, traj: Trajectory):
    """Differentiable U_n = traj(U_0)."""
    return _function().apply(U0, traj)


def density(plan, U):
    """Density of every cell of every block, a view of U (torch) in state order."""
    return U.view(-1, 4)[:, 0]


class AdjointProblem:
    """J(U_0) = sum over cells of (rho(U_n) - rho_target(U_n))^2 on one solver's plan, with the
    target trajectory computed once from a target initial state."""

    def __init__(self, solver, target_ic, nsteps: int = 100, checkpoint_every: int = 10):
        self.solver = solver
        self.plan = solver.plan
        P = self.plan
        torch = P.torch
        solver.apply_initial_condition()
        self.U0 = P.U.clone()
        self.dt = frozen_schedule(P, nsteps)
        self.traj = Trajectory(P, self.dt, checkpoint_every)
        # target: the same frozen schedule from the target initial state
        solver.apply_initial_condition(target_ic)
        self.U0_target = P.U.clone()
        self.traj.forward(self.U0_target)
        self.rho_target = density(P, P.U).clone()
        P.U.copy_(self.U0)
        self.torch = torch

    def objective(self, U0):
        UT = trajectory(U0, self.traj)
        r = density(self.plan, UT) - self.rho_target
        return (r * r).sum()

    def value(self, U0) -> float:
        """J without recording a graph (forward only)."""
        UT = self.traj.forward(U0)
        r = density(self.plan, UT) - self.rho_target
        return float((r * r).sum())

    def gradient(self, U0: Optional[object] = None):
        """(J, dJ/dU_0) at U0 (default: the base initial state)."""
        x = (self.U0 if U0 is None else U0).clone().requires_grad_(True)
        J = self.objective(x)
        J.backward()
        return float(J.detach()), x.grad.detach()
