"""Mutation check of the oracle's pins: each entry applies one plausible misreading of the reference
to the oracle (a textbook formula where the reference has a quirk, a dropped term, a wrong sign, a
swapped index) and asserts that the pin written against that mistake fails.  A pin that still
passes under its mutant pins nothing, so this file is what makes the pins in test_oracle_pins.py
meaningful.  CPU only; each case patches the oracle module for the duration of one test."""
import math

import numpy as np
import pytest

from oracle import pyhype_oracle as O

import test_oracle_pins as P


# ----------------------------------------------------------------------------- the misreadings
def roe_state_enthalpy(WL, WR):
    """Textbook Roe average: sqrt(rho)-weighted u, v and total enthalpy H, and p* recovered from H*
    (instead of the reference's sqrt(rho)-weighted pressure, states/primitive.py:301-320)."""
    sL, sR = np.sqrt(WL[..., 0]), np.sqrt(WR[..., 0])
    inv = 1.0 / (sL + sR)

    def H(W):
        return O.G_OVER_GM1 * W[..., 3] / W[..., 0] + 0.5 * (W[..., 1] ** 2 + W[..., 2] ** 2)

    Q = np.zeros(WL.shape)
    Q[..., 0] = np.sqrt(WL[..., 0] * WR[..., 0])
    Q[..., 1] = (WL[..., 1] * sL + WR[..., 1] * sR) * inv
    Q[..., 2] = (WL[..., 2] * sL + WR[..., 2] * sR) * inv
    Hs = (H(WL) * sL + H(WR) * sR) * inv
    Q[..., 3] = (Hs - 0.5 * (Q[..., 1] ** 2 + Q[..., 2] ** 2)) * Q[..., 0] / O.G_OVER_GM1
    return Q


def _patch_rc_uncorrected(mp):
    orig = O.roe_rc_data

    def rc(Wr, aroe, Lm, Lp):
        return orig(Wr, aroe, Wr[:, :, 1] - aroe, Wr[:, :, 1] + aroe)

    mp.setattr(O, "roe_rc_data", rc)


def _patch_lp_rho(mp):
    orig = O.roe_lp_data

    def lp(Wr, aroe):
        d = orig(Wr, aroe)
        num = Wr.shape[0] * Wr.shape[1]
        d[2 * num:3 * num] = np.ravel(Wr[:, :, 0])
        return d

    mp.setattr(O, "roe_lp_data", lp)


def _patch_harten_off(mp):
    mp.setattr(O, "harten", lambda R_p, R_m, L_p, L_m, Roe_p, Roe_m: (Roe_p.copy(), Roe_m.copy()))


def _patch_outer_own_limited(mp):
    orig = O._outer_state

    def outer(b, d, limited_boundary, ghost=None):
        if b.bc[d] is None and b.neighbors[d] is not None:
            return limited_boundary.copy()
        return orig(b, d, limited_boundary, ghost)

    mp.setattr(O, "_outer_state", outer)


def _patch_outer_neighbour_limited(mp):
    """Second-order interface: the neighbour's own limited face value at the shared face."""
    orig_outer, orig_bc = O._outer_state, O.apply_boundary_condition
    seen = {}

    def bc(blocks):
        seen["blocks"] = blocks
        return orig_bc(blocks)

    def outer(b, d, limited_boundary, ghost=None):
        if b.bc[d] is None and b.neighbors[d] is not None:
            nb = seen["blocks"][b.neighbors[d]]
            g = nb.geom
            gx, gy = O.green_gauss(nb.U, nb.ghost, g)
            hot = {k: O.high_order_term(gx, gy, g, k) for k in O.DIRS}
            phi = O.limiter_phi(nb.U, nb.ghost, hot)
            opp = O.OPPOSITE[d]
            face = nb.U + phi * hot[opp]
            return face[O._ghost_sl[opp]].copy()
        return orig_outer(b, d, limited_boundary, ghost)

    mp.setattr(O, "apply_boundary_condition", bc)
    mp.setattr(O, "_outer_state", outer)


def _nrm2_double(x):
    return np.sqrt(((x[..., 0] * x[..., 0] + x[..., 1] * x[..., 1]) + x[..., 2] * x[..., 2]) + x[..., 3] * x[..., 3])


def _patch_hlll_hll(mp):
    mp.setattr(O, "hlll_coef", lambda aroe, u, dU, dF, Lm, Lp: Lm * Lp)


def _patch_hlll_mx_min(mp):
    def coef(aroe, u, dU, dF, Lm, Lp):
        k = aroe * O.nrm2(dU)
        n2 = O.nrm2(dF - u[:, :, None] * dU)
        alpha = np.maximum(0, 1 - n2 / np.where(k < 1e-16, k + 1e-14, k))
        return Lm * Lp * (1 - alpha * (1 - np.minimum(u / Lm, u / Lp)))

    mp.setattr(O, "hlll_coef", coef)


def _patch_slope_default_zero(mp):
    orig = O.compute_slope

    def slope(dmax, dmin, d):
        s = orig(dmax, dmin, d)
        return np.where(d == 0, 0.0, s)

    mp.setattr(O, "compute_slope", slope)


def _patch_phi_clamped(mp):
    orig = O.venkatakrishnan
    mp.setattr(O, "venkatakrishnan", lambda s: np.minimum(1.0, orig(s)))


def _patch_reflect_tangential(mp):
    def reflect(state, wall_theta):
        O.rotate(wall_theta, state)
        state[:, :, 2] = -state[:, :, 2]
        O.unrotate(wall_theta, state)

    mp.setattr(O, "reflect", reflect)


def _patch_gg_no_area(mp):
    orig = O.green_gauss

    def gg(U, ghost, g):
        gx, gy = orig(U, ghost, g)
        return gx * g.A[..., None], gy * g.A[..., None]

    mp.setattr(O, "green_gauss", gg)


def _patch_ek_without_v(mp):
    def c2p(U):
        rho = U[:, :, 0]
        u = U[:, :, 1] / rho
        v = U[:, :, 2] / rho
        p = O.GM1 * (U[:, :, 3] - rho * (0.5 * (u * u)))
        return np.dstack((rho.copy(), u, v, p))

    mp.setattr(O, "cons_to_prim", c2p)


def _patch_energy_flux_without_p(mp):
    orig = O.prim_F

    def F(W):
        out = orig(W)
        out[:, :, 3] = W[:, :, 1] * (O.ONE_OVER_GM1 * W[:, :, 3] + O.prim_ek(W))
        return out

    mp.setattr(O, "prim_F", F)


def _patch_dt_swapped(mp):
    def dt(b, CFL):
        U = b.U
        rho = U[:, :, 0]
        u, v = U[:, :, 1] / rho, U[:, :, 2] / rho
        p = O.GM1 * (U[:, :, 3] - rho * (0.5 * (u * u + v * v)))
        a = np.sqrt(O.GAMMA * p / rho)
        return CFL * np.amin(np.minimum(b.geom.dy / (np.abs(u) + a), b.geom.dx / (np.abs(v) + a)))

    mp.setattr(O, "block_dt", dt)


def _patch_west_normal_flipped(mp):
    orig = O.block_geometry

    def geom(vert, nx, ny, nq=1):
        g = orig(vert, nx, ny, nq)
        g.face["W"].nx = -g.face["W"].nx
        return g

    mp.setattr(O, "block_geometry", geom)


def _patch_mesh_neighbours_swapped(mp):
    orig = O.rectangular_mesh_dict

    def mesh(*a, **k):
        m = orig(*a, **k)
        for d in m.values():
            d["NeighborE"], d["NeighborW"] = d["NeighborW"], d["NeighborE"]
        return m

    mp.setattr(O, "rectangular_mesh_dict", mesh)


def _patch_rk4_weights(mp):
    table = dict(O.BUTCHER)
    table["RK4"] = [[0.5], [0, 0.5], [0, 0, 1], [0.25, 0.25, 0.25, 0.25]]
    mp.setattr(O, "BUTCHER", table)


def _patch_gauss2_weights(mp):
    q = dict(O.QUADS)
    q[2] = ((-1 / np.sqrt(3), 0.5), (1 / np.sqrt(3), 0.5))
    mp.setattr(O, "QUADS", q)


def _run_rk_pins():
    with pytest.MonkeyPatch.context() as m:
        P.test_rk_stability_polynomial(m, "RK4", P._taylor(4), 1e-14)


# name -> (patch(monkeypatch), pins of which at least one must fail)
MUTANTS = {
    "B.1 enthalpy-based Roe pressure": (
        lambda mp: mp.setattr(O, "roe_state", roe_state_enthalpy),
        [P.test_roe_state_pressure_is_sqrt_rho_weighted, P.test_roe_mass_flux_uses_sqrt_rho_weighted_pressure]),
    "B.2 uncorrected speeds in Rc": (
        _patch_rc_uncorrected, [P.test_roe_symmetric_expansion_is_central]),
    "B.2/B.6 no Harten fix": (
        _patch_harten_off, [P.test_roe_symmetric_expansion_is_central, P.test_hlll_signal_speeds_carry_the_harten_fix]),
    "B.15 shear strength rho* dv": (
        _patch_lp_rho, [P.test_roe_shear_strength_is_dv]),
    "B.3 own limited state at block interfaces": (
        _patch_outer_own_limited, [P.test_interblock_face_takes_the_neighbour_cell_average]),
    "B.3 neighbour's limited state at block interfaces": (
        _patch_outer_neighbour_limited, [P.test_interblock_residual_ignores_the_neighbours_second_column]),
    "dnrm2 as a double-precision sum of squares": (
        lambda mp: mp.setattr(O, "nrm2", _nrm2_double), [P.test_nrm2_is_blas_dnrm2_bit_for_bit]),
    "HLLL without the alpha term (HLL)": (
        _patch_hlll_hll, [P.test_hlll_resolves_a_moving_contact_exactly]),
    "HLLL with min(u/S-, u/S+)": (
        _patch_hlll_mx_min, [P.test_hlll_resolves_a_moving_contact_exactly]),
    "B.4 limiter delta hot instead of (U + hot) - U": (
        lambda mp: mp.setattr(O, "reconstruction_delta", lambda U, h: h),
        [P.test_limiter_delta_is_reconstructed_value_minus_average]),
    "B.4 slope 0 (not 1) where the delta is zero": (
        _patch_slope_default_zero, [P.test_venkatakrishnan_closed_forms,
                                    P.test_limiter_on_one_dimensional_linear_field_is_three_quarters]),
    "B.4 limiter clamped to 1": (
        _patch_phi_clamped, [P.test_venkatakrishnan_closed_forms]),
    "reflection negates the tangential velocity": (
        _patch_reflect_tangential, [P.test_wall_faces_carry_no_mass_or_energy, P.test_ghost_fill_neighbours_and_walls]),
    "Green-Gauss without 1/A": (
        _patch_gg_no_area, [P.test_green_gauss_exact_on_linear_field]),
    "kinetic energy without v^2": (
        _patch_ek_without_v, [P.test_conservative_primitive_round_trip]),
    "energy flux without the p term": (
        _patch_energy_flux_without_p, [P.test_physical_flux_closed_form]),
    "CFL time step with dx and dy swapped": (
        _patch_dt_swapped, [P.test_dt_closed_form_on_uniform_state]),
    "west face normal with the wrong sign": (
        _patch_west_normal_flipped, [P.test_rectangle_geometry_closed_forms]),
    "east and west neighbours swapped": (
        _patch_mesh_neighbours_swapped, [P.test_multiblock_corner_interpolation, P.test_ghost_fill_neighbours_and_walls]),
    "RK4 final weights 1/4": (
        _patch_rk4_weights, [_run_rk_pins]),
    "2-point Gauss weights 1/2": (
        _patch_gauss2_weights, [lambda: P.test_gauss_rules_integrate_polynomials(2)]),
}


@pytest.mark.parametrize("name", list(MUTANTS))
def test_mutant_fails_its_pin(name, monkeypatch):
    patch, pins = MUTANTS[name]
    patch(monkeypatch)
    failed = []
    for pin in pins:
        try:
            with np.errstate(all="ignore"):
                pin()
        except Exception as e:  # AssertionError, or the mutant breaking the computation itself
            failed.append((getattr(pin, "__name__", "pin"), type(e).__name__))
    assert failed, f"no pin catches the mutant '{name}'"


def test_pins_pass_without_mutation():
    for _, pins in MUTANTS.values():
        for pin in pins:
            pin()
    assert math.isfinite(O.GM1)
