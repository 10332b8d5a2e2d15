"""phx_residual_vjp_noalloc: the residual VJP with its work arrays in the library's static device
array, called through the C ABI, against phx_residual_vjp (work arrays the plan allocates) bit
for bit -- on two calls in a row, the second on a smaller plan, so the second call starts from the
first call's leftovers in the shared array -- and the refusal of a plan too large for the array."""
import ctypes as C

import numpy as np
import pytest

from pyhype_amd import Euler2D
from pyhype_amd import _native as nat
from pyhype_amd import workloads as W

pytestmark = pytest.mark.gpu


def _setup(cfg, mesh, seed):
    import torch

    s = Euler2D(cfg, mesh)
    s.apply_initial_condition()
    rng = np.random.default_rng(seed)
    U = s.plan.U.clone()
    U *= 1.0 + 1e-3 * torch.from_numpy(rng.standard_normal(U.shape[0])).to(U.device)
    lamR = torch.from_numpy(rng.standard_normal(U.shape[0])).to(U.device)
    return s, U, lamR


def _vjp(s, U, lamR, fn):
    import torch

    P = s.plan
    lamU = torch.full_like(U, float("nan"))
    rc = getattr(P.lib, fn)(P.handle, P.ptr(U), P.ptr(lamR), P.ptr(lamU), P.stream())
    nat.check(rc, fn)
    torch.cuda.synchronize()
    return lamU.cpu().numpy()


def test_noalloc_vjp_equals_the_allocating_vjp_on_two_calls():
    # call 1: 2x1 blocks of 24x24 (rotated-frame geometry of the reference's rectangular generator), Roe
    cfg, mesh = W.explosion_multiblock(24, 2, 1)
    s1, U1, l1 = _setup(cfg, mesh, seed=3)
    print("no-alloc VJP: call 1, 2 blocks of 24x24, Roe", flush=True)
    got1 = _vjp(s1, U1, l1, "phx_residual_vjp_noalloc")
    ref1 = _vjp(s1, U1, l1, "phx_residual_vjp")
    assert np.isfinite(ref1).all()
    assert np.array_equal(got1, ref1)
    # call 2: a smaller plan (3 blocks of 10x10, HLLL, outlets and a wall): the shared array still
    # holds call 1's work arrays
    cfg2, mesh2 = W.oblique_shock(10, 3)
    s2, U2, l2 = _setup(cfg2, mesh2, seed=4)
    print("no-alloc VJP: call 2, 3 blocks of 10x10, HLLL", flush=True)
    got2 = _vjp(s2, U2, l2, "phx_residual_vjp_noalloc")
    ref2 = _vjp(s2, U2, l2, "phx_residual_vjp")
    assert np.isfinite(ref2).all()
    assert np.array_equal(got2, ref2)
    # and again on the first plan, after the second
    assert np.array_equal(_vjp(s1, U1, l1, "phx_residual_vjp_noalloc"), ref1)
    s1.plan.close()
    s2.plan.close()


def test_noalloc_vjp_refuses_a_plan_larger_than_its_array():
    # 400x400 cells need (16 + 20) * 8 B per cell + 64 B per face, about 67 MB > 32 MiB
    cfg, mesh = W.explosion_single(400)
    s = Euler2D(cfg, mesh)
    P = s.plan
    lam = P.U.clone()
    rc = P.lib.phx_residual_vjp_noalloc(P.handle, P.ptr(P.U), P.ptr(lam), P.ptr(lam), P.stream())
    assert rc == nat.PHX_E_INVALID
    msg = P.lib.phx_last_error().decode()
    assert "static array" in msg and str(32 * 1024 * 1024) in msg
    P.close()
