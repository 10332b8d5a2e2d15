"""The kernels' per-face / per-cell arithmetic (pyhype_amd/csrc/phx_numerics.hpp, built for the host)
against the oracle, bit for bit, on seeded random inputs including the branches the reference
takes (Harten fix active, HLLL supersonic branches, limiter d == 0, walls)."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyhype_oracle as O

PHYS = np.array([O.GAMMA, O.GM1, O.ONE_OVER_GM1, O.G_OVER_GM1, 0.8])


def dp(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))


def rand_prim(rng, n, umax=1.0):
    rho = rng.uniform(0.2, 5.0, n)
    u = rng.uniform(-umax, umax, n)
    v = rng.uniform(-umax, umax, n)
    p = rng.uniform(0.1, 5.0, n)
    return np.stack([rho, u, v, p], axis=1)


def face_pairs(rng, n):
    WL = rand_prim(rng, n, 2.0)
    WR = rand_prim(rng, n, 2.0)
    # nearly equal pairs (Harten correction active near sonic points) and identical pairs
    k = n // 4
    WR[:k] = WL[:k] * (1 + 1e-3 * rng.standard_normal((k, 4)))
    WL[k:2 * k, 1] = np.sqrt(O.GAMMA * WL[k:2 * k, 3] / WL[k:2 * k, 0]) * rng.uniform(0.95, 1.05, k)
    WR[2 * k:2 * k + 10] = WL[2 * k:2 * k + 10]
    # supersonic to the right and to the left (HLLL upwind branches)
    WL[-20:-10, 1] = 5.0
    WR[-20:-10, 1] = 5.0
    WL[-10:, 1] = -5.0
    WR[-10:, 1] = -5.0
    return WL, WR


@pytest.mark.parametrize("flux", ["Roe", "HLLL"])
def test_face_flux_bitwise(numerics_shim, flux):
    rng = np.random.default_rng(1234)
    n = 4000
    WL, WR = face_pairs(rng, n)
    F = np.zeros((n, 4))
    numerics_shim.shim_prim_flux(0 if flux == "Roe" else 1, n, dp(WL), dp(WR), dp(F), dp(PHYS))
    fn = O.flux_roe if flux == "Roe" else O.flux_hlll
    ref = fn(WL.reshape(1, n, 4), WR.reshape(1, n, 4)).reshape(n, 4)
    np.testing.assert_array_equal(F, ref)


def test_cons_to_prim_and_reflect(numerics_shim):
    rng = np.random.default_rng(7)
    n = 1000
    U = O.prim_to_cons(rand_prim(rng, n).reshape(1, n, 4)).reshape(n, 4)
    W = np.zeros((n, 4))
    numerics_shim.shim_cons_to_prim(n, dp(U), dp(W), dp(PHYS))
    np.testing.assert_array_equal(W, O.cons_to_prim(U.reshape(1, n, 4)).reshape(n, 4))
    th = rng.uniform(-np.pi, np.pi, n)
    th[:10] = np.pi / 2
    th[10:20] = 0.0
    c = np.array([np.cos(t) for t in th])
    import math
    c = np.array([math.cos(t) for t in th])
    s = np.array([math.sin(t) for t in th])
    out = np.zeros((n, 4))
    numerics_shim.shim_reflect(n, dp(U), dp(c), dp(s), dp(out))
    ref = U.reshape(n, 1, 4).copy()
    O.reflect(ref, th.reshape(n, 1))
    np.testing.assert_array_equal(out, ref.reshape(n, 4))


@pytest.mark.parametrize("lim,name", [(-1, "Venkatakrishnan"), (0, "Venkatakrishnan"), (1, "BarthJespersen"),
                                      (2, "VanAlbada"), (3, "VanLeer")])
def test_reconstruct_matches_oracle_on_block(numerics_shim, lim, name):
    """Cell reconstruction inside a random block with walls, using the oracle's geometry; every
    limiter of the reference (lim = -1: the default-compiled Venkatakrishnan path)."""
    rng = np.random.default_rng(99)
    nx, ny = 13, 9
    for vert in ({"SW": [0, 0], "SE": [10, 0], "NW": [0, 7], "NE": [10, 7]},
                 {"SW": [0, 0], "SE": [4.999999999999999, 0.1], "NW": [0.2, 5.000000000000001], "NE": [5, 5]}):
        g = O.block_geometry(vert, nx, ny)
        W = rand_prim(rng, nx * ny).reshape(ny, nx, 4)
        W[2:5, 3:6] = W[2, 3]          # a uniform patch (limiter d == 0 branch)
        U = O.prim_to_cons(W)
        b = O.Block(0, g, {d: None for d in O.DIRS}, {d: "Reflection" for d in O.DIRS}, U=U)
        O.apply_boundary_condition({0: b})
        gx, gy = O.green_gauss(U, b.ghost, g)
        hot = {d: O.high_order_term(gx, gy, g, d) for d in O.DIRS}
        phi = O.limiter_phi(U, b.ghost, hot, O.LIMITERS[name])
        ref = np.concatenate([U + phi * hot[d] for d in ("E", "W", "N", "S")], axis=2)
        EW = np.concatenate((b.ghost["W"], U, b.ghost["E"]), axis=1)
        NS = np.concatenate((b.ghost["S"], U, b.ghost["N"]), axis=0)
        Uw, Ue, Us, Un = EW[:, :-2], EW[:, 2:], NS[:-2], NS[2:]
        f = g.face
        cols = [f["E"].L * f["E"].nx, f["W"].L * f["W"].nx, f["N"].L * f["N"].nx, f["S"].L * f["S"].nx,
                f["E"].L * f["E"].ny, f["W"].L * f["W"].ny, f["N"].L * f["N"].ny, f["S"].L * f["S"].ny,
                1.0 / g.A,
                f["E"].qx - g.xc, f["W"].qx - g.xc, f["N"].qx - g.xc, f["S"].qx - g.xc,
                f["E"].qy - g.yc, f["W"].qy - g.yc, f["N"].qy - g.yc, f["S"].qy - g.yc]
        G = np.stack(cols, axis=-1).reshape(-1, 17)
        out = np.zeros((nx * ny, 16))
        if lim < 0:
            numerics_shim.shim_reconstruct(nx * ny, dp(U), dp(Uw), dp(Ue), dp(Us), dp(Un), dp(G), dp(out))
        else:
            numerics_shim.shim_reconstruct_lim(lim, nx * ny, dp(U), dp(Uw), dp(Ue), dp(Us), dp(Un), dp(G), dp(out))
        np.testing.assert_array_equal(out.reshape(ny, nx, 16), ref)


def test_cell_dt(numerics_shim):
    rng = np.random.default_rng(5)
    g = O.block_geometry({"SW": [0, 0], "SE": [3, 0], "NW": [0, 2], "NE": [3, 2]}, 11, 7)
    U = O.prim_to_cons(rand_prim(rng, 77).reshape(7, 11, 4))
    b = O.Block(0, g, {}, {}, U=U)
    out = np.zeros(77)
    numerics_shim.shim_cell_dt(77, dp(U), dp(g.dx), dp(g.dy), dp(out), dp(PHYS))
    assert 0.8 * out.min() == O.block_dt(b, 0.8)


def test_hlll_norm_is_blas_dnrm2_bit_for_bit(numerics_shim):
    """The kernels' nrm4 (64-bit-significand emulation of the x87 dnrm2 that Numba's np.linalg.norm
    reaches in flux/HLLL.py:86-93) equals scipy.linalg.blas.dnrm2 bit for bit on 10^6 seeded vectors
    over twelve decades, the dU / dF - u dU vectors of HLLL faces, exact squares and tie-prone integer
    vectors, subnormal and huge magnitudes (squares outside the double range), zeros and signed zeros."""
    from scipy.linalg import blas
    rng = np.random.default_rng(31)
    X = [rng.standard_normal((1_000_000, 4)) * 10.0 ** rng.uniform(-6, 6, (1_000_000, 1))]
    WL, WR = face_pairs(rng, 20000)
    UL = O.prim_to_cons(WL.reshape(1, -1, 4))[0]
    UR = O.prim_to_cons(WR.reshape(1, -1, 4))[0]
    dU = UR - UL
    X.append(dU)
    X.append((O.prim_F(WR.reshape(1, -1, 4)) - O.prim_F(WL.reshape(1, -1, 4)))[0] - rng.uniform(-2, 2, (20000, 1)) * dU)
    X.append(np.ldexp(rng.integers(1, 2 ** 30, (20000, 4)).astype(float), rng.integers(-40, 40, (20000, 1))))
    for scale in (1e-300, 1e-165, 1e-150, 1e150, 1e200, 1e300):
        X.append(rng.standard_normal((2000, 4)) * scale)
    X.append(np.array([[0.0, 0.0, 0.0, 0.0], [-0.0, 0.0, -0.0, 0.0], [3.0, 4.0, 0.0, 0.0], [0.0, 0.0, 0.0, 5e-324],
                       [1.7e308, 0.0, 0.0, 0.0], [1.0, 1.0, 1.0, 1.0]]))
    X = np.ascontiguousarray(np.vstack(X))
    ref = np.array([blas.dnrm2(x) for x in X])
    got = np.zeros(len(X))
    numerics_shim.shim_nrm4(len(X), dp(X), dp(got))
    bad = np.nonzero(got.view(np.int64) != ref.view(np.int64))[0]
    assert bad.size == 0, f"{bad.size} of {len(X)} differ, e.g. {X[bad[:3]]}"


def _near_midpoint_vectors(rng, n):
    """(v, w, t, 0) with v^2 + w^2 close to (v + ulp(v)/2)^2: norms within a few 64-bit units of a
    midpoint between two doubles, where the x87 sum's rounding to 64 bits decides the result; random
    signs and orders, a third component 2^-20..2^-60 below w on half of them."""
    v = np.ldexp(rng.uniform(1.0, 2.0, n), rng.integers(-100, 100, n))
    u = np.spacing(v)
    pert = np.where(rng.random(n) < 0.5, rng.uniform(-1e-6, 1e-6, n),
                    np.ldexp(rng.uniform(-1.0, 1.0, n), -rng.integers(0, 60, n)))
    w = np.sqrt(v * u + 0.25 * u * u) * (1.0 + pert)
    t = np.where(rng.random(n) < 0.5, 0.0, w * np.ldexp(rng.uniform(0.5, 1.0, n), -rng.integers(20, 60, n)))
    X = np.stack([v, w, t, np.zeros(n)], axis=1) * np.where(rng.random((n, 4)) < 0.5, -1.0, 1.0)
    swap = rng.random(n) < 0.5
    X[swap, 0], X[swap, 1] = X[swap, 1].copy(), X[swap, 0].copy()
    return X


def test_hlll_norm_paths_at_rounding_ties(numerics_shim):
    """nrm4 evaluates the extended sum in pairs of doubles and takes the integer emulation only where
    that cannot decide a rounding.  Both paths equal scipy.linalg.blas.dnrm2 bit for bit on the inputs
    that stress the decisions: norms at midpoints between doubles (double rounding, ties to even),
    sums of powers of two (exact 64-bit ties in the additions), short-mantissa integers, spreads of
    up to 2^160 between components and magnitudes at the fast path's range limits (2^-400, 2^400)."""
    from scipy.linalg import blas
    rng = np.random.default_rng(47)
    n = 60000
    X = [_near_midpoint_vectors(rng, n)]
    X.append(np.ldexp(np.where(rng.random((n, 4)) < 0.5, -1.0, 1.0), rng.integers(-30, 30, (n, 4))))
    X.append(np.ldexp(rng.integers(1, 2 ** 30, (n, 4)).astype(float), rng.integers(-20, 20, (n, 4))))
    X.append(np.where(rng.random((n, 4)) < 0.3, 0.0,
                      rng.uniform(-1, 1, (n, 4)) * np.exp2(rng.uniform(-80, 80, (n, 4)))))
    X.append(rng.uniform(-1, 1, (n, 4)) * np.exp2(rng.uniform(-420, 420, (n, 4))))
    X.append(np.array([[2.0 ** 400, 2.0 ** -400, 0.0, 0.0], [2.0 ** 401, 1.0, 0.0, 0.0],
                       [2.0 ** -400, 2.0 ** -401, 0.0, 0.0], [1.0, 2.0 ** -27, 0.0, 0.0],
                       [1.0, 2.0 ** -32, 2.0 ** -32, 0.0]]))
    # 1 + 2^-62 + b^2, b = 2^-32 (1 + k 2^-52): the sum is 2^-115 k above a 64-bit tie whose even
    # neighbour lies below, and the low parts span 54 bits, so the pair sum alone rounds the wrong way
    X.append(np.array([[1.0, 2.0 ** -31, 2.0 ** -32 * (1.0 + k * 2.0 ** -52), 0.0] for k in (1, 3, 5, 7)]))
    X = np.ascontiguousarray(np.vstack(X))
    ref = np.array([blas.dnrm2(x) for x in X])
    for name in ("shim_nrm4", "shim_nrm4_ext"):
        got = np.zeros(len(X))
        getattr(numerics_shim, name)(len(X), dp(X), dp(got))
        bad = np.nonzero(got.view(np.int64) != ref.view(np.int64))[0]
        assert bad.size == 0, f"{name}: {bad.size} of {len(X)} differ, e.g. {X[bad[:3]]}"


def test_hlll_norm_pair_sum_is_the_x87_sum(numerics_shim):
    """nrm4's fast path holds the x87 sum of squares (each product and sum rounded to a 64-bit
    significand) as a pair of doubles.  Where it reports the sum decided, hi + lo equals NumPy's
    80-bit long double evaluation of ((a^2 + b^2) + c^2) + d^2 exactly: on seeded vectors, on ties
    from powers of two and short integers, and on sums 2^-115 k above a 64-bit tie whose low parts
    span 54 bits (there the rounded pair sum alone would round to the wrong side).  The fast path
    decides at least 99.5% of the seeded vectors."""
    if np.finfo(np.longdouble).nmant != 63:
        pytest.skip("long double is not the x87 80-bit format on this host")
    rng = np.random.default_rng(53)
    n = 50000
    X = [rng.standard_normal((n, 4)) * 10.0 ** rng.uniform(-6, 6, (n, 1)),
         np.ldexp(np.where(rng.random((n, 4)) < 0.5, -1.0, 1.0), rng.integers(-30, 30, (n, 4))),
         np.ldexp(rng.integers(1, 2 ** 30, (n, 4)).astype(float), rng.integers(-20, 20, (n, 4))),
         np.array([[1.0, 2.0 ** -31, 2.0 ** -32 * (1.0 + k * 2.0 ** -52), 0.0] for k in (1, 3, 5, 7)])]
    X = np.ascontiguousarray(np.vstack(X))
    m = len(X)
    hi, lo, ok = np.zeros(m), np.zeros(m), np.zeros(m, dtype=np.int32)
    numerics_shim.shim_sum64(m, dp(X), dp(hi), dp(lo), ok.ctypes.data_as(C.POINTER(C.c_int)))
    L = X.astype(np.longdouble)
    ref = ((L[:, 0] * L[:, 0] + L[:, 1] * L[:, 1]) + L[:, 2] * L[:, 2]) + L[:, 3] * L[:, 3]
    got = hi.astype(np.longdouble) + lo.astype(np.longdouble)
    dec = ok == 1
    bad = np.nonzero(dec & (got != ref))[0]
    assert bad.size == 0, f"{bad.size} decided sums differ, e.g. {X[bad[:3]]}"
    assert dec[:n].mean() >= 0.995
    assert not dec[-4:].any()   # the 54-bit low-part cases are left to the emulation


def test_roe_flat_faces_bitwise(numerics_shim):
    """Faces whose two states give jumps of +0 take roe_flux's flat-region shortcut,
    F = 0.5*(F_L + F_R): bit for bit (signs of zero included) the oracle's full Roe formula, on
    equal states with zero, negative and supersonic velocities, both signs of zero on either side,
    and densities and pressures at and beyond the shortcut's range limits (2^-300, 2^300)."""
    rng = np.random.default_rng(99)
    rows = []
    for u in (0.0, -0.0, 0.3, -0.7, 5.0, -5.0, 1e-300, -1e-300):
        for v in (0.0, -0.0, 0.2, -3.0):
            for rho, p in ((1.0, 1.0), (0.3, 7.0), (2.0 ** -299, 2.0 ** -299), (2.0 ** 299, 2.0 ** 299),
                           (2.0 ** -301, 1.0), (1.0, 2.0 ** 301), (1.7, 1e-12)):
                rows.append(((rho, u, v, p), (rho, u, v, p)))
    # a +0 on one side and -0 on the other: the jump is +0 (flat) or -0 (not flat)
    for a, b in ((0.0, -0.0), (-0.0, 0.0)):
        rows.append(((1.2, a, 0.1, 0.9), (1.2, b, 0.1, 0.9)))
        rows.append(((1.2, 0.4, a, 0.9), (1.2, 0.4, b, 0.9)))
    W = rand_prim(rng, 200, 3.0)
    rows += [(tuple(w), tuple(w)) for w in W]
    WL = np.array([r[0] for r in rows], dtype=np.float64)
    WR = np.array([r[1] for r in rows], dtype=np.float64)
    n = len(rows)
    F = np.zeros((n, 4))
    numerics_shim.shim_prim_flux(0, n, dp(WL), dp(WR), dp(F), dp(PHYS))
    with np.errstate(all="ignore"):
        ref = O.flux_roe(WL.reshape(1, n, 4), WR.reshape(1, n, 4)).reshape(n, 4)
    assert np.array_equal(F.view(np.int64), ref.view(np.int64))


def test_venkatakrishnan_face_minimum_equals_the_np_minimum_chain(numerics_shim):
    """The four faces' limiter values reduce with np.minimum in the order E, W, N, S
    (limiters/base.py:47-187).  The Venkatakrishnan path takes fmin where no value is NaN, which is
    exact because its values are never -0; here the values come from the limiter itself (random
    slopes, slopes near -2 and 0, the flat-face 3/4, overflowing slopes that give NaN), and the
    result must equal the chain and numpy's reduction bit for bit."""
    rng = np.random.default_rng(11)
    n = 200_000
    d = rng.standard_normal(n) * np.exp(rng.uniform(-40, 40, n))
    d[rng.random(n) < 0.1] = 0.0                              # flat faces: 3/4
    dmax = np.abs(rng.standard_normal(n)) * np.exp(rng.uniform(-40, 40, n))
    dmin = -np.abs(rng.standard_normal(n)) * np.exp(rng.uniform(-40, 40, n))
    dmax[rng.random(n) < 0.1] = 0.0                           # local extrema: s = +-0
    dmin[rng.random(n) < 0.1] = -0.0
    k = n // 20
    dmin[:k] = -2.0 * np.abs(d[:k])                           # s = -2 exactly where d > 0 is not taken
    dmax[k:2 * k] = 1e200
    d[k:2 * k] = 1e-200                                       # s overflows: s*s = inf -> NaN
    vals = np.empty(n)
    numerics_shim.shim_venkat_face(n, dp(dmax), dp(dmin), dp(d), dp(vals))
    assert np.isnan(vals).any() and (vals == 0.75).any() and (vals == 0.0).any()
    assert not np.any((vals == 0.0) & np.signbit(vals)), "a Venkatakrishnan value was -0"
    faces = vals[rng.integers(0, n, (n, 4))]
    fast, chain = np.empty(n), np.empty(n)
    numerics_shim.shim_limiter_min4(n, dp(faces), dp(fast), dp(chain))
    ref = np.minimum(np.minimum(np.minimum(faces[:, 0], faces[:, 1]), faces[:, 2]), faces[:, 3])
    assert np.isnan(ref).any()
    assert np.array_equal(fast.view(np.uint64), chain.view(np.uint64))
    same = (fast.view(np.uint64) == ref.view(np.uint64)) | (np.isnan(fast) & np.isnan(ref))
    assert same.all()
