"""Tangent-linear (forward-mode) products on the GPU (gpu), and the dot-product test of the
discrete adjoint.

* phx_residual_jvp equals the host build of the same element operations (whose value part the CPU
  tests check bit for bit against the residual and whose derivative against central differences);
* phx_step_jvp equals the Runge-Kutta recurrence composed from residual JVPs;
* Trajectory.jvp agrees with central differences of the trajectory on a short run;
* <J v, w> = <v, J^T w> for the 256^2 adjoint workload (100 SSP-RK2 steps with its frozen dt
  schedule, checkpoints every 10) on 16 seeded direction pairs: the reverse sweep is the exact
  transpose of the independently evaluated forward-mode derivative at full size.
"""
import ctypes as C

import numpy as np
import pytest

from pyhype_amd import Euler2D
from pyhype_amd import workloads as W
from pyhype_amd.adjoint import AdjointProblem, Trajectory, frozen_schedule
from tests.test_adjoint_cpu import CASES, MIDPOINT, _p, _phys, _smooth_state, dp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def shim(adjoint_shim):
    adjoint_shim.shim_residual_jvp.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, dp, dp, dp, dp]
    return adjoint_shim


def _solver(name):
    (cfg, mesh), general = CASES[name]()
    return cfg, Euler2D(cfg, mesh, force_general=general)


@pytest.mark.parametrize("name", MIDPOINT)
def test_residual_jvp_matches_host_build(shim, name):
    import torch

    cfg, s = _solver(name)
    P = s.plan
    U = _smooth_state(P, seed=4, mach=0.9 if cfg.fvm_flux_function_type == "HLLL" else 0.4)
    v = np.random.default_rng(3).standard_normal(U.size) * np.abs(U)
    print(f"residual JVP: {name}", flush=True)
    got = P.residual_jvp(torch.as_tensor(U, device="cuda"), torch.as_tensor(v, device="cuda")).cpu().numpy()
    Rv, ref = np.zeros_like(U), np.zeros_like(U)
    flux = 0 if cfg.fvm_flux_function_type == "Roe" else 1
    assert shim.shim_residual_jvp(P._descs, len(P.blocks), flux, _p(_phys(cfg)), _p(U), _p(v), _p(Rv), _p(ref)) == 0
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    print(name, "bitwise equal:", np.array_equal(got, ref))


def test_residual_jvp_refuses_gauss_faces():
    import torch

    cfg, mesh = W.explosion_single(8)
    cfg.fvm_num_quadrature_points = 2
    P = Euler2D(cfg, mesh).plan
    with pytest.raises(ValueError, match="midpoint faces only"):
        P.residual_jvp(P.U, torch.zeros_like(P.U))


@pytest.mark.parametrize("integrator", ["SSPRK2", "RK4"])
def test_step_jvp_is_the_rk_recurrence(integrator):
    import torch

    cfg, mesh = W.explosion_multiblock(12, 2, 2)
    cfg.time_integrator = integrator
    P = Euler2D(cfg, mesh).plan
    U0 = torch.as_tensor(_smooth_state(P, seed=5), device="cuda")
    P.U.copy_(U0)
    dt = P.dt()
    stages = torch.zeros((1, P.nstages, P.state_len), dtype=torch.float64, device="cuda")
    P.run_record(dt.view(1, -1), stages)
    dU0 = torch.as_tensor(np.random.default_rng(6).standard_normal(P.state_len), device="cuda") * U0.abs()
    dU = dU0.clone()
    P.step_jvp(stages[0], dt, dU)
    # reference: dU_i = dU_0 + dt sum_{m<i} a[i-1][m] dK_m, dK_m = JVP(U_m, dU_m)
    a, S, d = P._a, P.nstages, float(dt[0])
    dK, dUi = [], dU0
    for i in range(S):
        dK.append(P.residual_jvp(stages[0, i].contiguous(), dUi.contiguous()))
        dUi = dU0 + sum(d * a[i, m] * dK[m] for m in range(i + 1) if a[i, m] != 0.0)
    assert torch.allclose(dU, dUi, rtol=1e-12, atol=1e-12 * float(dUi.abs().max()))


def test_trajectory_jvp_matches_central_differences():
    """12x12 smoothed explosion, 4 SSP-RK2 steps with a frozen schedule, 4 seeded directions."""
    import torch

    cfg, mesh, target, _ = W.adjoint_problem(12, seed=0)
    s = Euler2D(cfg, mesh)
    s.apply_initial_condition()
    P = s.plan
    U0 = P.U.clone()
    dt = frozen_schedule(P, 4)
    traj = Trajectory(P, dt, checkpoint_every=2)
    rng = np.random.default_rng(9)
    for trial in range(4):
        v = torch.as_tensor(rng.standard_normal(P.state_len), device="cuda") * U0.abs()
        jv = traj.jvp(U0, v)
        h = 1e-7
        up = traj.forward(U0 + h * v)
        um = traj.forward(U0 - h * v)
        fd = (up - um) / (2 * h)
        err = float((fd - jv).norm() / jv.norm())
        print(f"trajectory JVP vs central differences, direction {trial}: {err:.2e}", flush=True)
        assert err < 1e-6


@pytest.mark.slow
def test_dot_product_of_tangent_and_adjoint_at_full_size():
    """<J v, w> against <v, J^T w> on the 256^2 adjoint workload (100 SSP-RK2 steps, frozen dt
    schedule, checkpoints every 10), 16 seeded pairs of unit-norm Gaussian directions: the
    difference within 1e-12 of |Jv||w| (and 1e-9 of the value)."""
    import torch

    cfg, mesh, target, _ = W.adjoint_problem(256, seed=0)
    prob = AdjointProblem(Euler2D(cfg, mesh), target, nsteps=100, checkpoint_every=10)
    traj, U0 = prob.traj, prob.U0
    rng = np.random.default_rng(2024)
    worst = 0.0
    for trial in range(16):
        v = torch.as_tensor(rng.standard_normal(U0.numel()), device="cuda")
        w = torch.as_tensor(rng.standard_normal(U0.numel()), device="cuda")
        v /= v.norm()
        w /= w.norm()
        jv = traj.jvp(U0, v)
        traj.forward(U0)
        jtw = traj.backward(w)
        lhs, rhs = float(jv @ w), float(v @ jtw)
        # rounding of a dot product of n terms is relative to |Jv||w| (Cauchy-Schwarz), not to the
        # (much smaller, for random directions) value itself
        scale = max(float(jv.norm()), float(jtw.norm()))
        rel = abs(lhs - rhs) / scale
        worst = max(worst, rel)
        print(f"pair {trial}: <Jv,w> = {lhs:.15e}, <v,J^T w> = {rhs:.15e}, |diff|/(|Jv||w|) {rel:.2e}, "
              f"|diff|/|<Jv,w>| {abs(lhs - rhs) / abs(lhs):.2e}", flush=True)
        assert abs(lhs - rhs) <= 1e-9 * abs(lhs)
    assert worst < 1e-12, worst
