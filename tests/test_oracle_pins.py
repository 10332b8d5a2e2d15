"""Pins of the CPU oracle by mathematics (no reference import is possible here: numba and mpi4py
are absent).  Each check is chosen so that a plausible misreading -- a dropped term, a wrong sign
or index, a transposed operand, a wrong COO entry -- fails it."""
import math

import numpy as np
import pytest

from oracle import pyhype_oracle as O

RECT = {"SW": [0.0, 0.0], "SE": [10.0, 0.0], "NW": [0.0, 5.0], "NE": [10.0, 5.0]}


def mesh1(bc="Reflection", vert=RECT):
    d = {"nBLK": 0, **vert}
    for s in O.DIRS:
        d["Neighbor" + s] = None
        d["BCType" + s] = bc
    return {0: d}


# ------------------------------------------------------------------ geometry
def test_rectangle_geometry_closed_forms():
    g = O.block_geometry(RECT, 20, 10)
    assert g.is_cartesian
    np.testing.assert_allclose(g.A, 0.5 * 0.5, rtol=1e-13)
    np.testing.assert_allclose(g.face["E"].L, 0.5, rtol=1e-14)
    np.testing.assert_allclose(g.face["N"].L, 0.5, rtol=1e-14)
    # normals: E (1, 0), W (-1, -0), N (cos(pi/2), 1), S (-cos(pi/2), -1)
    assert np.all(g.face["E"].nx == 1.0) and np.all(g.face["E"].ny == 0.0)
    assert np.all(g.face["W"].nx == -1.0)
    assert np.all(g.face["N"].nx == math.cos(math.pi / 2)) and np.all(g.face["N"].ny == 1.0)
    assert np.all(g.face["S"].ny == -1.0)
    np.testing.assert_allclose(g.xc[0], np.linspace(0.25, 9.75, 20), rtol=1e-14)
    np.testing.assert_allclose(g.dx, 0.5, rtol=1e-14)
    # each face's quadrature point is its midpoint
    for d in O.DIRS:
        np.testing.assert_allclose(g.face[d].qx, g.face[d].mx, rtol=1e-15)


def test_skewed_geometry_area_and_normals():
    vert = {"SW": [0.0, 0.0], "SE": [4.0, 1.0], "NW": [1.0, 3.0], "NE": [5.0, 4.0]}   # parallelogram
    g = O.block_geometry(vert, 8, 6)
    assert not g.is_cartesian
    total = (4.0 * 3.0 - 1.0 * 1.0)
    np.testing.assert_allclose(g.A.sum(), total, rtol=1e-12)
    for d in O.DIRS:   # unit normals
        np.testing.assert_allclose(g.face[d].nx ** 2 + g.face[d].ny ** 2, 1.0, rtol=1e-14)
    # closed cell: sum of L*n over the four faces vanishes
    sx = sum(g.face[d].L * g.face[d].nx for d in O.DIRS)
    sy = sum(g.face[d].L * g.face[d].ny for d in O.DIRS)
    assert np.max(np.abs(sx)) < 1e-13 and np.max(np.abs(sy)) < 1e-13


def test_multiblock_corner_interpolation():
    """The rectangular generator's transfinite interior corners are off by ulps -- the reason the
    2x4 multiblock mesh runs the reference's rotated-frame path."""
    m = O.rectangular_mesh_dict(["Reflection"] * 1, ["Reflection"], ["Reflection"], ["Reflection"],
                                10.0, 0.0, 20.0, 0.0, 2, 4)
    ne = m[0]["NE"]
    assert abs(ne[0] - 5.0) < 1e-14 and abs(ne[1] - 5.0) < 1e-14
    assert not (ne[0] == 5.0 and ne[1] == 5.0)
    assert m[0]["NeighborE"] == 1 and m[0]["NeighborN"] == 2 and m[7]["NeighborS"] == 5
    assert m[0]["BCTypeE"] is None and m[1]["BCTypeE"] == "Reflection"


# ------------------------------------------------------------------ gradient and limiter
def _block_with(U, vert=RECT, bc="OutletDirichlet"):
    ny, nx, _ = U.shape
    g = O.block_geometry(vert, nx, ny)
    b = O.Block(0, g, {d: None for d in O.DIRS}, {d: bc for d in O.DIRS}, U=U)
    O.apply_boundary_condition({0: b})
    return b


def test_green_gauss_exact_on_linear_field():
    nx, ny = 12, 8
    g = O.block_geometry(RECT, nx, ny)
    U = np.stack([2.0 + 0.3 * g.xc - 0.7 * g.yc, 1 + 0 * g.xc, 0.5 * g.yc, 3 + g.xc], axis=-1)
    b = _block_with(U)
    gx, gy = O.green_gauss(U, b.ghost, g)
    inner = np.s_[1:-1, 1:-1]
    np.testing.assert_allclose(gx[inner][..., 0], 0.3, rtol=1e-12)
    np.testing.assert_allclose(gy[inner][..., 0], -0.7, rtol=1e-12)
    np.testing.assert_allclose(gy[inner][..., 2], 0.5, rtol=1e-12)
    np.testing.assert_allclose(gx[inner][..., 3], 1.0, rtol=1e-12)
    assert np.max(np.abs(gx[inner][..., 1])) < 1e-14


def test_venkatakrishnan_closed_forms():
    s = np.array([0.0, 1.0, 2.0, 1e12])
    np.testing.assert_allclose(O.venkatakrishnan(s), [0.0, 0.75, 1.0, 1.0], rtol=1e-11)
    # zero slope denominators default to s = 1, i.e. phi = 0.75; no upper clamp
    assert O.compute_slope(np.array([1.0]), np.array([-1.0]), np.array([0.0]))[0] == 1.0
    smax = 2 + 2 * math.sqrt(2)          # d(phi)/ds = 0: phi_max = 1.0938...
    assert O.venkatakrishnan(np.array([smax]))[0] > 1.09
    grid = np.linspace(0, 50, 20001)
    assert np.max(O.venkatakrishnan(grid)) <= O.venkatakrishnan(np.array([smax]))[0] + 1e-15


def test_other_limiter_closed_forms():
    """BarthJespersen min(1, s); VanAlbada (s^2+s)/(s^2+1); VanLeer (|s|+s)/(s+1) = 2s/(s+1) on
    s >= 0 (limiters/limiters.py:25-62): phi(0) = 0, phi(1) = 1 for all three, symmetric limiters
    satisfy phi(s)/s = phi(1/s), and the limits at s -> infinity are 1, 1 and 2."""
    s = np.array([0.0, 0.5, 1.0, 2.0, 1e12])
    np.testing.assert_array_equal(O.barth_jespersen(s), [0.0, 0.5, 1.0, 1.0, 1.0])
    np.testing.assert_allclose(O.van_albada(s), [0.0, 0.6, 1.0, 1.2, 1.0], rtol=1e-11)
    np.testing.assert_allclose(O.van_leer(s), [0.0, 2 / 3, 1.0, 4 / 3, 2.0], rtol=1e-11)
    r = np.linspace(0.1, 10, 50)
    for fn in (O.van_albada, O.van_leer):
        np.testing.assert_allclose(fn(r) / r, fn(1 / r), rtol=1e-12)
    assert np.isnan(O.barth_jespersen(np.array([np.nan])))[0]          # np.minimum propagates NaN


@pytest.mark.parametrize("name,value", [("Venkatakrishnan", 0.75), ("BarthJespersen", 1.0), ("VanAlbada", 1.0),
                                        ("VanLeer", 1.0)])
def test_limiter_on_one_dimensional_linear_field(name, value):
    """x-linear field: E/W faces give s = 2, N/S faces have zero reconstruction delta (s = 1):
    phi = min(f(2), f(1)) in every interior cell -- 0.75 for Venkatakrishnan, 1 for the others."""
    nx, ny = 10, 6
    g = O.block_geometry(RECT, nx, ny)
    U = np.stack([1.0 + 0.1 * g.xc, 1 + 0 * g.xc, 1 + 0 * g.xc, 5 + 0.2 * g.xc], axis=-1)
    b = _block_with(U)
    gx, gy = O.green_gauss(U, b.ghost, g)
    hot = {d: O.high_order_term(gx, gy, g, d) for d in O.DIRS}
    phi = O.limiter_phi(U, b.ghost, hot, O.LIMITERS[name])
    inner = np.s_[1:-1, 1:-1]
    np.testing.assert_allclose(phi[inner][..., 0], value, rtol=1e-14)
    np.testing.assert_allclose(phi[inner][..., 3], value, rtol=1e-14)


def test_limiter_on_one_dimensional_linear_field_is_three_quarters():
    """x-linear field: E/W faces give s = 2 (phi = 1), N/S faces have zero reconstruction delta
    (s = 1, phi = 0.75): the reference limits every interior cell of a 1-D ramp to 0.75."""
    nx, ny = 10, 6
    g = O.block_geometry(RECT, nx, ny)
    U = np.stack([1.0 + 0.1 * g.xc, 1 + 0 * g.xc, 1 + 0 * g.xc, 5 + 0.2 * g.xc], axis=-1)
    b = _block_with(U)
    gx, gy = O.green_gauss(U, b.ghost, g)
    hot = {d: O.high_order_term(gx, gy, g, d) for d in O.DIRS}
    phi = O.limiter_phi(U, b.ghost, hot)
    inner = np.s_[1:-1, 1:-1]
    np.testing.assert_array_equal(phi[inner][..., 0], 0.75)
    np.testing.assert_array_equal(phi[inner][..., 3], 0.75)


def test_limiter_delta_is_reconstructed_value_minus_average():
    """limiters/base.py:96-107 with fvm/SecondOrderMUSCL.py:86-94: the slope's delta is
    (U + hot) - U.  On a uniform state with hot = 1e-20 (far below half an ulp of U = 1) that delta
    is exactly 0, so s = 1 and phi = 0.75 -- a delta taken as hot itself would give s = dmax/hot = 0
    and phi = 0."""
    U = np.ones((3, 4, 4))
    ghost = {"E": U[:, -1:].copy(), "W": U[:, :1].copy(), "N": U[-1:].copy(), "S": U[:1].copy()}
    hot = {d: np.full(U.shape, 1e-20) for d in O.DIRS}
    np.testing.assert_array_equal(O.limiter_phi(U, ghost, hot), 0.75)


# ------------------------------------------------------------------ fluxes
def test_physical_flux_closed_form():
    """PrimitiveState._F_from_prim_JIT (states/primitive.py:223-237): W = (2, 0.5, -1.5, 3) gives
    F = (rho u, rho u^2 + p, rho u v, u (p/(gamma-1) + rho (u^2+v^2)/2 + p)) = (1, 3.5, -1.5, 6.5)."""
    F = O.prim_F(_face(2.0, 0.5, -1.5, 3.0))[0, 0]
    np.testing.assert_allclose(F, [1.0, 3.5, -1.5, 6.5], rtol=1e-15)


def test_conservative_primitive_round_trip():
    """states/converter/concrete_defs.py:85-141: W = (2, 0.5, -1.5, 3) -> U = (2, 1, -3,
    3/(gamma-1) + 2.5) -> W again, with both velocity components in the kinetic energy."""
    U = O.prim_to_cons(_face(2.0, 0.5, -1.5, 3.0))
    np.testing.assert_allclose(U[0, 0], [2.0, 1.0, -3.0, 3.0 / 0.4 + 2.5], rtol=1e-15)
    np.testing.assert_allclose(O.cons_to_prim(U)[0, 0], [2.0, 0.5, -1.5, 3.0], rtol=1e-14)


def _prim(rng, n, umax=1.0):
    return np.stack([rng.uniform(0.5, 3, n), rng.uniform(-umax, umax, n), rng.uniform(-umax, umax, n),
                     rng.uniform(0.5, 3, n)], axis=1).reshape(1, n, 4)


@pytest.mark.parametrize("fn", [O.flux_roe, O.flux_hlll])
def test_flux_consistency(fn):
    W = _prim(np.random.default_rng(0), 200, 2.0)
    F = fn(W, W.copy())
    np.testing.assert_allclose(F, O.prim_F(W), rtol=1e-13, atol=1e-13)


def test_roe_eigen_layout_inverts():
    """The COO blocks of Rc (with uncorrected eigenvalues) and Lp satisfy Lp * dW/dU * Rc =
    diag(1, 1, 1/rho*, 1) at the Roe state: the reference never sets the (2,2) entry of Lp (it keeps
    the 1.0 of np.ones, flux/Roe.py:222-251, flux/eigen_system.py:52-57), so its shear-wave strength
    is dv instead of rho* dv.  Any other misplaced entry or sign breaks the identity."""
    rng = np.random.default_rng(1)
    F = 5
    WL, WR = _prim(rng, F), _prim(rng, F)
    Wr = O.roe_state(WL, WR)
    a = O.prim_a(Wr)
    u, v = Wr[0, :, 1], Wr[0, :, 2]
    Ek = 0.5 * (u * u + v * v)
    H = O.G_OVER_GM1 * Wr[0, :, 3] / Wr[0, :, 0] + Ek
    Rc_i, Rc_j, Lp_i, Lp_j, _ = O._roe_coo(F)
    ua = u * a[0]
    Rc_data = np.ones(14 * F)
    for k, val in ((0, H - ua), (1, v), (2, Ek), (3, u - a[0]), (4, v), (5, v), (7, u), (9, H + ua), (11, v),
                   (12, u + a[0])):
        Rc_data[k * F:(k + 1) * F] = val
    ia = 1.0 / a[0]
    r2a = 0.5 * Wr[0, :, 0] * ia
    Lp_data = np.ones(7 * F)
    for k, val in ((0, r2a), (3, 0.5 * ia * ia), (4, -r2a), (5, -ia * ia), (6, 0.5 * ia * ia)):
        Lp_data[k * F:(k + 1) * F] = val
    Rc = np.zeros((4 * F, 4 * F))
    Lp = np.zeros((4 * F, 4 * F))
    np.add.at(Rc, (Rc_i, Rc_j), Rc_data)
    np.add.at(Lp, (Lp_i, Lp_j), Lp_data)
    for f in range(F):
        sl = slice(4 * f, 4 * f + 4)
        rho = Wr[0, f, 0]
        uu, vv = u[f], v[f]
        # dW/dU for W = (rho, u, v, p), U = (rho, rho u, rho v, E)
        M = np.array([[1, 0, 0, 0],
                      [-uu / rho, 1 / rho, 0, 0],
                      [-vv / rho, 0, 1 / rho, 0],
                      [O.GM1 * Ek[f], -O.GM1 * uu, -O.GM1 * vv, O.GM1]])
        np.testing.assert_allclose(Lp[sl, sl] @ M @ Rc[sl, sl], np.diag([1.0, 1.0, 1.0 / rho, 1.0]), atol=1e-12)


def test_hlll_supersonic_branches_are_upwind():
    W = _prim(np.random.default_rng(2), 50)
    WL, WR = W.copy(), W[:, ::-1].copy()
    WL[..., 1] = WR[..., 1] = 6.0
    np.testing.assert_array_equal(O.flux_hlll(WL, WR), O.flux_hlll(WL, WL))
    WL[..., 1] = WR[..., 1] = -6.0
    np.testing.assert_array_equal(O.flux_hlll(WL, WR), O.flux_hlll(WR, WR))


def test_wall_faces_carry_no_mass_or_energy():
    """A state against its own reflection: zero mass and energy flux for Roe always, and for HLLL
    while the wall-normal speed is below ~0.85 a.  Above that (flow leaving the wall), the Harten
    fix applied to HLLL's signal speeds (flux/HLLL.py:33-37) makes S+ != -S- and the wall face
    carries mass -- a property of the reference scheme, asserted here too."""
    rng = np.random.default_rng(3)
    W = _prim(rng, 40)
    a = np.sqrt(O.GAMMA * W[..., 3] / W[..., 0])
    W[..., 1] = a * rng.uniform(-0.5, 0.5, a.shape)
    U = O.prim_to_cons(W)
    Ur = U.copy()
    O.reflect(Ur, np.zeros((1, 40)))
    assert np.array_equal(Ur[..., 1], -U[..., 1]) and np.array_equal(Ur[..., 2], U[..., 2])
    for fn in (O.flux_roe, O.flux_hlll):
        F = fn(O.cons_to_prim(U), O.cons_to_prim(Ur))
        assert np.max(np.abs(F[..., 0])) < 1e-14 and np.max(np.abs(F[..., 3])) < 1e-13
    W[..., 1] = -0.95 * a                      # strong outflow from the wall
    U = O.prim_to_cons(W)
    Ur = U.copy()
    O.reflect(Ur, np.zeros((1, 40)))
    assert np.max(np.abs(O.flux_roe(O.cons_to_prim(U), O.cons_to_prim(Ur))[..., 0])) < 1e-14
    assert np.min(np.abs(O.flux_hlll(O.cons_to_prim(U), O.cons_to_prim(Ur))[..., 0])) > 1e-6


# ------------------------------------------------------------------ residual and time stepping
@pytest.mark.parametrize("bc", ["Reflection", "OutletDirichlet"])
def test_rest_state_is_steady(bc):
    sim = O.Simulation(mesh1(bc), 9, 7, CFL=0.8)
    U = O.prim_to_cons(np.tile(np.array([1.3, 0.0, 0.0, 2.0]), (7, 9, 1)))
    sim.set_state({0: U})
    assert np.max(np.abs(O.residual(sim.blocks[0], sim.flux_fn))) == 0.0


def test_uniform_flow_with_outlets_is_steady():
    sim = O.Simulation(mesh1("OutletDirichlet"), 9, 7, CFL=0.8)
    U = O.prim_to_cons(np.tile(np.array([1.3, 0.4, -0.2, 2.0]), (7, 9, 1)))
    sim.set_state({0: U})
    assert np.max(np.abs(O.residual(sim.blocks[0], sim.flux_fn))) == 0.0


@pytest.mark.parametrize("flux", ["Roe", "HLLL"])
def test_single_block_walls_conserve_mass_and_energy(flux):
    sim = O.Simulation(mesh1("Reflection"), 16, 12, flux=flux, CFL=0.5, a_inf=343.0)
    U = O.explosion_state(sim.blocks[0].geom, box=(2.0, 6.0, 1.0, 3.0))
    sim.set_state({0: U})
    A = sim.blocks[0].geom.A[..., None]
    m0 = (U * A).sum(axis=(0, 1))
    sim.run(6)
    m1 = (sim.states()[0] * A).sum(axis=(0, 1))
    assert abs(m1[0] - m0[0]) / m0[0] < 1e-13
    assert abs(m1[3] - m0[3]) / m0[3] < 1e-13


def test_residual_mirror_equivariance_on_smooth_state():
    """R(mirror U) = mirror R(U) for a smooth generic state (x- and y-mirrors and the transpose).
    On piecewise-constant data the reference arithmetic is not mirror-exact: the cos(pi/2) leakage
    of the N/S normals rounds mirrored cells differently by an ulp, and the limiter's d > 0 / d == 0
    branches turn that into O(1) changes of phi -- so the check uses data without exact ties."""
    n = 16
    vert = {"SW": [0.0, 0.0], "SE": [10.0, 0.0], "NW": [0.0, 10.0], "NE": [10.0, 10.0]}
    sim = O.Simulation(mesh1("Reflection", vert), n, n, CFL=0.8)
    g = sim.blocks[0].geom
    x, y = g.xc, g.yc
    bump = lambda cx, cy: np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / 4.0)
    rho = 1.0 + 0.3 * (bump(5, 5) + 0.2 * bump(2.5, 5) + 0.2 * bump(7.5, 5) + 0.1 * bump(5, 2.5) + 0.1 * bump(5, 7.5))
    p = 1.0 + 0.5 * rho ** 2
    W = np.stack([rho, 0 * x, 0 * x, p], axis=-1)
    U = O.prim_to_cons(W)
    sim.set_state({0: U})
    R = O.residual(sim.blocks[0], sim.flux_fn)
    sc = np.abs(R).max(axis=(0, 1)) + 1e-300
    Mx = lambda A: A[:, ::-1] * np.array([1, -1, 1, 1])
    sim.set_state({0: Mx(U)})
    Rm = O.residual(sim.blocks[0], sim.flux_fn)
    assert np.max(np.abs(Rm - Mx(R)) / sc) < 1e-10
    T = lambda A: A.transpose(1, 0, 2)[..., [0, 2, 1, 3]]
    sim.set_state({0: T(U)})
    Rt = O.residual(sim.blocks[0], sim.flux_fn)
    assert np.max(np.abs(Rt - T(R)) / sc) < 1e-10


def _taylor(n, extra=()):
    return lambda z: sum(z ** k / math.factorial(k) for k in range(n + 1)) + sum(c * z ** k for k, c in extra)


@pytest.mark.parametrize("name,poly,rtol", [
    ("ExplicitEuler1", _taylor(1), 1e-14),
    ("RK2", _taylor(2), 1e-14),
    ("SSPRK2", _taylor(2), 1e-14),
    ("Ralston2", _taylor(2), 1e-14),
    ("RK3", _taylor(3), 1e-14),
    ("RK3SSP", _taylor(3), 1e-14),
    ("Ralston3", _taylor(3), 1e-14),
    ("RK4", _taylor(4), 1e-14),
    ("Ralston4", _taylor(4), 1e-7),          # the reference's 8-digit coefficients
    # the 6-stage 5th-order Dormand-Prince weights: R(z) = exp(z) to z^5, then z^6/600
    ("DormandPrince5", _taylor(5, [(6, 1 / 600)]), 1e-14),
    ("Generic2(0.3)", _taylor(2), 1e-14),
    ("Generic3(0.4)", _taylor(3), 1e-14),
])
def test_rk_stability_polynomial(monkeypatch, name, poly, rtol):
    """A linear residual R(U) = lam*U makes one step multiply U by the method's stability polynomial:
    exp(z) up to the method's order (plus the known z^6 term of Dormand-Prince), z = lam*dt."""
    if name.startswith("Generic"):
        table = (O.generic2 if name[7] == "2" else O.generic3)(float(name[9:-1]))
    else:
        table = O.BUTCHER[name]
    monkeypatch.setattr(O, "residual", lambda b, f: lam * b.U)
    monkeypatch.setattr(O, "apply_boundary_condition", lambda blocks: None)
    for lam, dt in ((-0.37, 0.9), (-1.3, 0.7), (0.6, 0.5)):
        g = O.block_geometry(RECT, 3, 2)
        b = O.Block(0, g, {}, {}, U=np.full((2, 3, 4), 2.0))
        O.integrate({0: b}, dt, table, None)
        np.testing.assert_allclose(b.U, 2.0 * poly(lam * dt), rtol=rtol)


def test_dt_closed_form_on_uniform_state():
    """QuadBlock.get_dt (blocks/quad_block.py:423-436): CFL min(dx / (|u| + a), dy / (|v| + a)) on
    cells of 0.5 x 0.25 (dx != dy, so a swap of the two is caught)."""
    sim = O.Simulation(mesh1(), 20, 20, CFL=0.8)
    U = O.prim_to_cons(np.tile(np.array([1.4, 0.3, -0.1, 1.0]), (20, 20, 1)))
    sim.set_state({0: U})
    a = math.sqrt(1.4 * 1.0 / 1.4)
    np.testing.assert_allclose(sim.get_dt(), 0.8 * min(0.5 / (0.3 + a), 0.25 / (0.1 + a)), rtol=1e-14)


def test_ghost_fill_neighbours_and_walls():
    m = O.rectangular_mesh_dict(["Reflection"], ["OutletDirichlet"], ["Reflection"], ["Reflection"],
                                4.0, 0.0, 2.0, 0.0, 2, 1)
    sim = O.Simulation(m, 5, 4)
    rng = np.random.default_rng(9)
    st = {n: O.prim_to_cons(_prim(rng, 20).reshape(4, 5, 4)) for n in m}
    sim.set_state(st)
    b0, b1 = sim.blocks[0], sim.blocks[1]
    np.testing.assert_array_equal(b0.ghost["E"], st[1][:, :1])     # neighbour's first column
    np.testing.assert_array_equal(b1.ghost["W"], st[0][:, -1:])    # neighbour's last column
    np.testing.assert_array_equal(b0.ghost["W"], st[0][:, :1])     # outlet: own column
    g = b1.ghost["E"]
    assert np.array_equal(g[..., 1], -st[1][:, -1:, 1]) and np.array_equal(g[..., 3], st[1][:, -1:, 3])


# ------------------------------------------------------------------ Gauss faces (2 and 3 points)
@pytest.mark.parametrize("nq", [1, 2, 3])
def test_gauss_rules_integrate_polynomials(nq):
    """An nq-point Gauss-Legendre rule integrates x^k over [-1, 1] exactly for k <= 2nq - 1 and
    not for k = 2nq; its weights sum to 2."""
    pts = [(float(p), float(w)) for p, w in O.QUADS[nq]]
    for k in range(2 * nq + 1):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        quad = sum(w * p ** k for p, w in pts)
        if k <= 2 * nq - 1:
            assert abs(quad - exact) < 1e-15
        else:
            assert abs(quad - exact) > 1e-3
    assert [p for p, _ in pts] == sorted(p for p, _ in pts)


def test_gauss_points_lie_on_faces_in_rule_order():
    """Points of a face lie on the segment between its two vertices, at the mapped Gauss abscissae,
    the first point nearest the first vertex (E: NE, W: NW, N: NE, S: SE); the points of a face
    shared by two cells coincide (E of cell j with W of cell j+1, N of row k with S of row k+1)."""
    vert = {"SW": [0.0, 0.0], "SE": [4.0, 1.0], "NW": [1.0, 3.0], "NE": [5.0, 4.0]}
    g = O.block_geometry(vert, 6, 5, 3)
    X, Y = g.xn, g.yn
    first = {"E": (X[1:, 1:], Y[1:, 1:]), "W": (X[1:, :-1], Y[1:, :-1]), "N": (X[1:, 1:], Y[1:, 1:]),
             "S": (X[:-1, 1:], Y[:-1, 1:])}
    second = {"E": (X[:-1, 1:], Y[:-1, 1:]), "W": (X[:-1, :-1], Y[:-1, :-1]), "N": (X[1:, :-1], Y[1:, :-1]),
              "S": (X[:-1, :-1], Y[:-1, :-1])}
    for d in O.DIRS:
        (ax, ay), (bx, by) = first[d], second[d]
        for (qx, qy, _), t in zip(g.qpts[d], (-math.sqrt(0.6), 0.0, math.sqrt(0.6))):
            s = 0.5 * (t + 1.0)     # fraction of the way from the first to the second vertex
            np.testing.assert_allclose(qx, ax + s * (bx - ax), atol=1e-14)
            np.testing.assert_allclose(qy, ay + s * (by - ay), atol=1e-14)
    for (ex, ey, _), (wx, wy, _) in zip(g.qpts["E"], g.qpts["W"]):
        assert np.array_equal(ex[:, :-1], wx[:, 1:]) and np.array_equal(ey[:, :-1], wy[:, 1:])
    for (nx_, ny_, _), (sx, sy, _) in zip(g.qpts["N"], g.qpts["S"]):
        assert np.array_equal(nx_[:-1], sx[1:]) and np.array_equal(ny_[:-1], sy[1:])


def _linear_field(x, y):
    U = np.empty(x.shape + (4,))
    U[..., 0] = 1.0 + 0.3 * x + 0.2 * y
    U[..., 1] = 0.4 + 0.25 * x - 0.1 * y
    U[..., 2] = -0.2 + 0.15 * x + 0.3 * y
    U[..., 3] = 2.5 + 0.5 * x + 0.35 * y
    return U


def _exact_residual(g, m=10):
    """-(1/A) times the boundary integral of F(U).n of the linear field, m-point Gauss-Legendre
    per edge (numpy's leggauss): the divergence the scheme approximates."""
    p, w = np.polynomial.legendre.leggauss(m)
    X, Y = g.xn, g.yn
    ne, nw, se, sw = (X[1:, 1:], Y[1:, 1:]), (X[1:, :-1], Y[1:, :-1]), (X[:-1, 1:], Y[:-1, 1:]), (X[:-1, :-1], Y[:-1, :-1])
    tot = 0.0
    for a, b in ((sw, se), (se, ne), (ne, nw), (nw, sw)):      # counter-clockwise, normal (dy, -dx)
        dx, dy = b[0] - a[0], b[1] - a[1]
        for pi, wi in zip(p, w):
            t = 0.5 * (pi + 1.0)
            U = _linear_field(a[0] + t * dx, a[1] + t * dy)
            r = U[..., 0]
            u, v = U[..., 1] / r, U[..., 2] / r
            pr = 0.4 * (U[..., 3] - 0.5 * r * (u * u + v * v))
            Fx = np.stack([r * u, r * u * u + pr, r * u * v, u * (U[..., 3] + pr)], -1)
            Fy = np.stack([r * v, r * u * v, r * v * v + pr, v * (U[..., 3] + pr)], -1)
            tot = tot + 0.5 * wi * (Fx * dy[..., None] - Fy * dx[..., None])
    return -tot / g.A[..., None]


def _gauss_error(n, nq, vert):
    d = {"nBLK": 0, **vert}
    for s in O.DIRS:
        d["Neighbor" + s] = None
        d["BCType" + s] = "OutletDirichlet"
    sim = O.Simulation({0: d}, n, n, limiter="BarthJespersen", quadrature_points=nq)
    b = sim.blocks[0]
    sim.set_state({0: _linear_field(b.geom.xc, b.geom.yc)})
    err = np.abs(O.residual(b, O.flux_roe) - _exact_residual(b.geom))
    return err[2:-2, 2:-2].max()


@pytest.mark.parametrize("vert", [{"SW": [0, 0], "SE": [1, 0], "NW": [0, 1], "NE": [1, 1]},
                                  {"SW": [0, 0], "SE": [1, 0.2], "NW": [0.3, 1], "NE": [1.3, 1.2]}],
                         ids=["rectangle", "parallelogram"])
def test_gauss_faces_converge_at_their_order(vert):
    """On a linear field the reconstruction is exact away from the boundary (Green-Gauss is exact on
    these meshes and BarthJespersen is inactive), the Roe flux of equal states is the physical flux,
    so the residual error is the face quadrature's alone: order 2 for the midpoint rule, 4 for
    2 points, round-off for 3 points at these sizes.  A weight permuted or a point placed on the
    wrong face fails these bounds."""
    e1 = [_gauss_error(n, 1, vert) for n in (8, 16)]
    e2 = [_gauss_error(n, 2, vert) for n in (8, 16)]
    e3 = _gauss_error(8, 3, vert)
    assert math.log2(e1[0] / e1[1]) > 1.4
    assert math.log2(e2[0] / e2[1]) > 3.2
    assert e2[0] < 1e-3 * e1[0]
    assert e3 < 1e-11


@pytest.mark.parametrize("nq", [2, 3])
def test_gauss_limiter_bounds_every_point(nq):
    """BarthJespersen with phi = min over all Gauss points of all faces keeps the limited value at
    every point inside the 5-point stencil's [min, max] (a limiter that looked at one point per face
    would let the others overshoot on rough data)."""
    rng = np.random.default_rng(7)
    vert = {"SW": [0.0, 0.0], "SE": [4.0, 1.0], "NW": [1.0, 3.0], "NE": [5.0, 4.0]}
    g = O.block_geometry(vert, 9, 7, nq)
    U = 1.0 + rng.random((7, 9, 4))
    b = O.Block(0, g, {s: None for s in O.DIRS}, {s: "OutletDirichlet" for s in O.DIRS}, U=U)
    O.apply_boundary_condition({0: b})
    gx, gy = O.green_gauss(U, b.ghost, g)
    hot = {d: [O.high_order_term(gx, gy, g, d, i) for i in range(nq)] for d in O.DIRS}
    phi = O.limiter_phi(U, b.ghost, hot, O.barth_jespersen)
    EW = np.concatenate((b.ghost["W"], U, b.ghost["E"]), axis=1)
    NS = np.concatenate((b.ghost["S"], U, b.ghost["N"]), axis=0)
    vals = (U, EW[:, :-2], EW[:, 2:], NS[:-2], NS[2:])
    hi, lo = np.maximum.reduce(vals), np.minimum.reduce(vals)
    limited_somewhere = False
    for d in O.DIRS:
        for h in hot[d]:
            q = U + phi * h
            assert np.all(q <= hi + 1e-13) and np.all(q >= lo - 1e-13)
            limited_somewhere |= bool(np.any((U + h > hi + 1e-9) | (U + h < lo - 1e-9)))
    assert limited_somewhere


@pytest.mark.parametrize("nq", [2, 3])
def test_gauss_face_states_pair_by_point(nq):
    """On a linear field the two limited states that meet at a Gauss point of an interior face are
    equal (both are the field's value there): point i of a cell's E face meets point i of its east
    neighbour's W face, likewise N with S.  Pairing point i with another point would differ by
    O(h * gradient)."""
    vert = {"SW": [0, 0], "SE": [1, 0.2], "NW": [0.3, 1], "NE": [1.3, 1.2]}
    g = O.block_geometry(vert, 10, 10, nq)
    U = _linear_field(g.xc, g.yc)
    b = O.Block(0, g, {s: None for s in O.DIRS}, {s: "OutletDirichlet" for s in O.DIRS}, U=U)
    O.apply_boundary_condition({0: b})
    gx, gy = O.green_gauss(U, b.ghost, g)
    hot = {d: [O.high_order_term(gx, gy, g, d, i) for i in range(nq)] for d in O.DIRS}
    phi = O.limiter_phi(U, b.ghost, hot, O.barth_jespersen)
    st = {d: [U + phi * h for h in hot[d]] for d in O.DIRS}
    inner = np.s_[2:-2, 2:-3]
    for i in range(nq):
        assert np.max(np.abs(st["E"][i][:, :-1] - st["W"][i][:, 1:])[inner]) < 1e-13
        assert np.max(np.abs(st["N"][i][:-1] - st["S"][i][1:])[2:-3, 2:-2]) < 1e-13
        if nq == 2:
            assert np.max(np.abs(st["E"][i][:, :-1] - st["W"][1 - i][:, 1:])[inner]) > 1e-3


# ------------------------------------------------------------------ reference quirks, closed forms
# Each expected value below is derived by hand from the cited reference lines (not by running the
# oracle's own formula); tests/test_oracle_mutations.py applies the textbook misreading of each
# quirk to the oracle and asserts that the pin written for it fails.
def _face(*w):
    return np.array(w, dtype=float).reshape(1, 1, 4)


def test_roe_state_pressure_is_sqrt_rho_weighted():
    """B.1, states/primitive.py:301-320: p* = (pL sqrt(rhoL) + pR sqrt(rhoR)) / (sqrt(rhoL) + sqrt(rhoR)).
    With rhoL = rhoR = 4 (sqrt = 2, weights 1/4) that is (pL + pR)/2 = 2 exactly, and u*, v* are the
    plain means.  The enthalpy-based (textbook) Roe pressure would add (gamma-1)/gamma rho du^2/8."""
    Q = O.roe_state(_face(4.0, 0.0, 0.3, 1.0), _face(4.0, 0.8, -0.2, 3.0))[0, 0]
    assert Q[0] == 4.0 and Q[1] == 0.4 and Q[3] == 2.0
    assert abs(Q[2] - 0.05) < 1e-16


def test_roe_mass_flux_uses_sqrt_rho_weighted_pressure():
    """B.1 through the flux: WL = (4, 0, 0, 1), WR = (4, 0.1, 0, 1).  dW = (0, 0.1, 0, 0), so only the
    acoustic strengths -+ rho* du / (2 a*) are non-zero, no Harten fix applies (|u* -+ a*| > 0.2),
    and the mass flux is 0.5 (0 + 0.4) - 0.5 r2a du (|u*+a*| - |u*-a*|) = 0.2 - 0.01 / a* with
    u* = 0.05, p* = 1, a* = sqrt(1.4 * 1 / 4)."""
    F = O.flux_roe(_face(4.0, 0.0, 0.0, 1.0), _face(4.0, 0.1, 0.0, 1.0))[0, 0]
    a = math.sqrt(1.4 / 4.0)
    assert abs(F[0] - (0.2 - 0.01 / a)) < 1e-15


def test_roe_symmetric_expansion_is_central():
    """B.2, flux/Roe.py:175-220,292-304: Rc carries the Harten-corrected speeds.  Symmetric
    expansion WL = (1.4, -0.5, 0, 1), WR = (1.4, 0.5, 0, 1), a = 1: u* = v* = 0, both speeds are
    corrected to the same lambda_c = 0.5 (a*^2/2 + 2) (|-+a*| < 2 (R - L) = 2), the strengths are
    -+ r2a, so Rc |Lambda| Lp dW = lambda_c r2a ((-1)(1, lambda_c, 0, H) + (1, lambda_c, 0, H)) = 0
    exactly: the flux is the central one, (0, rho w^2 + p, 0, 0).  With the uncorrected u -+ a in Rc
    the x-momentum would lose lambda_c a* r2a du."""
    F = O.flux_roe(_face(1.4, -0.5, 0.0, 1.0), _face(1.4, 0.5, 0.0, 1.0))[0, 0]
    assert F[0] == 0.0 and F[2] == 0.0 and F[3] == 0.0
    assert F[1] == (1.4 * 0.5) * 0.5 + 1.0


def test_roe_shear_strength_is_dv():
    """B.15, flux/Roe.py:222-251 with flux/eigen_system.py:52-57: the (2,2) entry of Lp keeps the 1.0
    of np.ones, so a pure shear jump (rho 4, u 0.5, p 1; v 0.25 | 1.25) has strength dv = 1 (not
    rho* dv = 4).  Only the shear column (0, 0, 1, v*) of Rc is active, |lambda| = u* = 0.5,
    v* = 0.75: the dissipation is 0.5 * 0.5 * (0, 0, 1, 0.75) = (0, 0, 0.25, 0.1875)."""
    WL, WR = _face(4.0, 0.5, 0.25, 1.0), _face(4.0, 0.5, 1.25, 1.0)
    F = O.flux_roe(WL, WR)[0, 0]
    central = 0.5 * (O.prim_F(WL) + O.prim_F(WR))[0, 0]
    np.testing.assert_allclose(central - F, [0.0, 0.0, 0.25, 0.1875], rtol=0, atol=2e-15)
    assert F[2] == 1.25      # 0.5 (0.5 + 2.5) - 0.25


def test_hlll_signal_speeds_carry_the_harten_fix():
    """B.6, flux/HLLL.py:27-37 with flux/base.py:119-146: symmetric expansion of speed w = 0.9 a
    (rho 1.4, p 1, a = 1).  The Roe speeds -+1 are corrected to lambda_c = 1/(8 w) + 2 w
    (= 1.93889 > w + a), so S+ = max(w + a, lambda_c) = lambda_c while S- = min(-w - a,
    +lambda_c) = -1.9.  alpha = max(0, 1 - |dF| / (a |dU|)) = 0 and u* = 0, so the flux is HLL:
    mass = rho w (w + a - lambda_c) / (lambda_c + w + a) = -0.0127637..., not 0."""
    w = 0.9
    lam = 1.0 / (8 * w) + 2 * w
    F = O.flux_hlll(_face(1.4, -w, 0.0, 1.0), _face(1.4, w, 0.0, 1.0))[0, 0]
    expect = 1.4 * w * (w + 1.0 - lam) / (lam + w + 1.0)
    assert abs(F[0] - expect) < 1e-11 * abs(expect)     # (w + a - lambda_c) cancels 2 digits


def test_hlll_resolves_a_moving_contact_exactly():
    """flux/HLLL.py:84-101: across a contact (rho 1 | 2, u 0.3, v 0.1, p 1) F(U) = u U + (0, p, 0, u p),
    so dF = u dU and |dF - u* dU| ~ 0: alpha = 1 and the coefficient is S- S+ max(u/S-, u/S+) =
    S- u, which makes the HLLL flux equal the upwind flux F(WL) (u > 0)."""
    WL, WR = _face(1.0, 0.3, 0.1, 1.0), _face(2.0, 0.3, 0.1, 1.0)
    np.testing.assert_allclose(O.flux_hlll(WL, WR), O.prim_F(WL), rtol=1e-13, atol=1e-15)


def test_nrm2_is_blas_dnrm2_bit_for_bit():
    """flux/HLLL.py:86-93: Numba's np.linalg.norm is the BLAS dnrm2 of its SciPy.  The oracle's nrm2
    equals scipy.linalg.blas.dnrm2 bit for bit on seeded vectors of HLLL magnitudes, on the dU and
    dF - u dU vectors of random faces, and on exact-square and tie-prone integer vectors."""
    from scipy.linalg import blas
    rng = np.random.default_rng(2024)
    X = [rng.standard_normal((40000, 4)) * 10.0 ** rng.uniform(-6, 6, (40000, 1))]
    X.append(np.ldexp(rng.integers(1, 2 ** 26, (5000, 4)).astype(float), rng.integers(-20, 20, (5000, 1))))
    X.append(np.array([[3.0, 4.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0], [1.0, 1.0, 1.0, 1.0], [-0.0, 2.0, 0.0, 0.0]]))
    W = _prim(rng, 5000, 2.0)
    WR = _prim(rng, 5000, 2.0)
    UL, UR = O.prim_to_cons(W), O.prim_to_cons(WR)
    dU = (UR - UL)[0]
    X.append(dU)
    X.append((O.prim_F(WR) - O.prim_F(W))[0] - rng.uniform(-1, 1, (5000, 1)) * dU)
    X = np.vstack(X)
    ref = np.array([blas.dnrm2(x) for x in X])
    got = O.nrm2(X)
    assert np.array_equal(got.view(np.int64), ref.view(np.int64))


def _two_blocks(nx=6, ny=4):
    """Blocks 0 | 1 on [0,5]x[0,4] and [5,10]x[0,4], joined at x = 5, walls elsewhere."""
    walls = "Reflection"
    def blk(n, x0, nb):
        d = {"nBLK": n, "SW": [x0, 0.0], "SE": [x0 + 5.0, 0.0], "NW": [x0, 4.0], "NE": [x0 + 5.0, 4.0]}
        for s in O.DIRS:
            d["Neighbor" + s] = nb.get(s)
            d["BCType" + s] = None if s in nb else walls
        return d
    return {0: blk(0, 0.0, {"E": 1}), 1: blk(1, 5.0, {"W": 0})}


def test_interblock_face_takes_the_neighbour_cell_average():
    """B.3, fvm/base.py:305-325: at a block interface the outer state is the neighbour's adjacent
    cell average (first order).  Block 0 at rest in W0 = (1, 0, 0, 1), block 1 at rest in
    W1 = (2, 0, 0, 3): block 0's last-column cells have phi = 0 (the west face's slope is
    dmin/d = 0 or dmax/d = 0), so their limited east value is W0, and their residual is
    L (F(W0) - Roe(W0, W1)) / A (the N/S fluxes cancel), everything else in block 0 is at rest."""
    m = _two_blocks()
    sim = O.Simulation(m, 6, 4)
    W0 = np.tile([1.0, 0.0, 0.0, 1.0], (4, 6, 1))
    W1 = np.tile([2.0, 0.0, 0.0, 3.0], (4, 6, 1))
    sim.set_state({0: O.prim_to_cons(W0), 1: O.prim_to_cons(W1)})
    R = O.residual(sim.blocks[0], sim.flux_fn)
    g = sim.blocks[0].geom
    Fc = O.prim_F(W0[:1, :1])[0, 0]
    Fr = O.flux_roe(W0[:1, :1].copy(), W1[:1, :1].copy())[0, 0]
    expect = g.face["E"].L[0, -1] * (Fc - Fr) / g.A[0, -1]
    for i in range(4):
        np.testing.assert_allclose(R[i, -1], expect, rtol=1e-13, atol=1e-12 * np.abs(expect).max())
    assert np.max(np.abs(R[:, :-1])) < 1e-12 * np.abs(expect).max()
    assert abs(expect[0]) > 0.1 and abs(expect[1]) > 0.1


def test_interblock_residual_ignores_the_neighbours_second_column():
    """B.3: block 0 sees block 1 only through block 1's first column (ghost cells and the outer
    state), so changing block 1's second column leaves block 0's residual bit-identical, while a
    change of its first column does not (a second-order outer state would read the second column
    through block 1's gradient)."""
    m = _two_blocks()
    rng = np.random.default_rng(11)
    sim = O.Simulation(m, 6, 4)
    st = {n: O.prim_to_cons(_prim(rng, 24).reshape(4, 6, 4)) for n in m}
    sim.set_state(st)
    R = O.residual(sim.blocks[0], sim.flux_fn)
    st2 = {n: v.copy() for n, v in st.items()}
    st2[1][:, 1] *= 1.05
    sim.set_state(st2)
    assert np.array_equal(O.residual(sim.blocks[0], sim.flux_fn), R)
    st3 = {n: v.copy() for n, v in st.items()}
    st3[1][:, 0] *= 1.05
    sim.set_state(st3)
    assert not np.array_equal(O.residual(sim.blocks[0], sim.flux_fn), R)
