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


# ------------------------------------------------------------------ fluxes
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
    sim = O.Simulation(mesh1(), 20, 10, CFL=0.8)
    U = O.prim_to_cons(np.tile(np.array([1.4, 0.3, -0.1, 1.0]), (10, 20, 1)))
    sim.set_state({0: U})
    a = math.sqrt(1.4 * 1.0 / 1.4)
    np.testing.assert_allclose(sim.get_dt(), 0.8 * min(0.5 / (0.3 + a), 0.5 / (0.1 + a)), rtol=1e-14)


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
