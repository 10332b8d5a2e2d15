"""CPU oracle for pyHype's second-order MUSCL residual and explicit Runge-Kutta path.

TEST INFRASTRUCTURE ONLY.  This module is a plain NumPy restatement of the
reference algorithm, written to check the HIP backend.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker (or the timed CPU baseline) -- never as the
thing measured or shipped.  The ``pyhype_amd`` package never imports it.

The restatement follows the reference's own array formulation: per-block
face arrays, ``np.concatenate`` of ghost strips, whole-sweep flux calls and,
for the Roe dissipation, real SciPy COO mat-vecs built with the reference's
block ordering, so that floating-point summation order is reproduced by
construction.  Every function cites the reference file:line it restates
(paths relative to the reference's ``pyhype/`` package).

Pinning (what checks this module, see tests/test_oracle_*.py):
  * geometry: closed forms for axis-aligned blocks, angle/normal identities;
  * Green-Gauss: exact on linear fields in interior cells;
  * limiter: closed-form Venkatakrishnan values, the s = 1 default, 1-D brute force;
  * Roe/HLLL: consistency F(W,W)=F(W), mirror symmetry, supersonic branches,
    an independent dense-matrix restatement of the eigen-decomposition, and
    hand-derived closed forms of the reference's own arithmetic: the
    sqrt(rho)-weighted Roe pressure (states/primitive.py:301-320), the
    Harten-corrected eigenvalues inside Rc (flux/Roe.py:175-220), the shear
    strength dv of the unset Lp(2,2) entry (flux/Roe.py:222-251);
  * HLLL norm: bit for bit equal to scipy.linalg.blas.dnrm2 (flux/HLLL.py:86-93);
  * inter-block faces: the neighbour's cell average as outer state
    (fvm/base.py:305-325), through a residual that ignores the neighbour's
    second column;
  * residual: rest-state invariance, conservation with reflecting walls;
  * RK: stability polynomials on a linear residual;
  * a mutation test (tests/test_oracle_mutations.py) applies plausible
    misreadings of these functions to this module and asserts that a pin fails
    for each;
  * end-to-end: golden fixtures in tests/golden/ (regression digests written
    by this module; they pin nothing about the reference by themselves).
Every function on the path has such a check; one without would be marked
"parity unpinned" here.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
from scipy.sparse import coo_matrix

# --------------------------------------------------------------------------------------
# Fluid constants  (fluids/base.py:58-64, fluids/air.py:27-28)
# --------------------------------------------------------------------------------------
GAMMA = 1.4
GM1 = GAMMA - 1              # (gamma - 1) as written in states/conservative.py:110 -> 0.3999999999999999
ONE_OVER_GM1 = 1.0 / (GAMMA - 1.0)   # fluids/base.py:62-64
G_OVER_GM1 = GAMMA / (GAMMA - 1.0)   # fluids/base.py:58-60

DIRS = ("E", "W", "N", "S")
OPPOSITE = {"E": "W", "W": "E", "N": "S", "S": "N"}


# --------------------------------------------------------------------------------------
# Mesh generation (mesh/base.py:133-226, mesh/rectangular.py:23-62)
# --------------------------------------------------------------------------------------
def rectangular_mesh_dict(BCE, BCW, BCN, BCS, east, west, north, south, nbh, nbv) -> Dict[int, dict]:
    """Block dictionaries of RectagularMeshGenerator.generate (mesh/rectangular.py:23-62),
    including the transfinite interpolation of interior block corners
    (mesh/base.py:88-121) and the block numbering/neighbour rules (mesh/base.py:191-226)."""
    nxc, nyc = nbh + 1, nbv + 1
    tb_x = np.linspace(west, east, nbh + 1)
    tb_y = np.ones(nbh + 1)
    ew_x = np.ones(nbv + 1)
    ew_y = np.linspace(south, north, nbv + 1)
    x, y = np.meshgrid(np.linspace(0, 1, nxc), np.linspace(0, 1, nyc))
    x[0, :] = tb_x
    y[0, :] = south * tb_y
    x[-1, :] = tb_x
    y[-1, :] = north * tb_y
    x[:, 0] = west * ew_x
    y[:, 0] = ew_y
    x[:, -1] = east * ew_x
    y[:, -1] = ew_y
    jm, im = np.meshgrid(np.linspace(1 / nxc, (nxc - 1) / nxc, nxc - 2),
                         np.linspace(1 / nyc, (nyc - 1) / nyc, nyc - 2))
    mj, mi = np.meshgrid(np.linspace((nxc - 1) / nxc, 1 / nxc, nxc - 2),
                         np.linspace((nyc - 1) / nyc, 1 / nyc, nyc - 2))

    def tfi(a):
        return (mi * a[0, 1:-1] + im * a[-1, 1:-1] + mj * a[1:-1, 0, None] + jm * a[1:-1, -1, None]
                - mi * mj * a[0, 0] - mi * jm * a[0, -1] - im * mj * a[-1, 0] - im * jm * a[-1, -1])

    x[1:-1, 1:-1] = tfi(x)
    y[1:-1, 1:-1] = tfi(y)
    BCE = [BCE[0]] * nbv if len(BCE) == 1 else BCE
    BCW = [BCW[0]] * nbv if len(BCW) == 1 else BCW
    BCN = [BCN[0]] * nbh if len(BCN) == 1 else BCN
    BCS = [BCS[0]] * nbh if len(BCS) == 1 else BCS
    out = {}
    for i in range(nbv):
        for j in range(nbh):
            n = nbh * i + j
            out[n] = {
                "nBLK": n,
                "NW": [x[i + 1, j], y[i + 1, j]], "NE": [x[i + 1, j + 1], y[i + 1, j + 1]],
                "SW": [x[i, j], y[i, j]], "SE": [x[i, j + 1], y[i, j + 1]],
                "NeighborE": n + 1 if j < nbh - 1 else None,
                "NeighborW": n - 1 if j > 0 else None,
                "NeighborN": n + nbh if n + nbh < nbv * nbh else None,
                "NeighborS": n - nbh if n - nbh >= 0 else None,
                "BCTypeE": BCE[i] if j == nbh - 1 else None,
                "BCTypeW": BCW[i] if j == 0 else None,
                "BCTypeN": BCN[j] if i == nbv - 1 else None,
                "BCTypeS": BCS[j] if i == 0 else None,
            }
    return out


# --------------------------------------------------------------------------------------
# Block geometry (mesh/quad_mesh.py:71-212, mesh/base.py:29-78, mesh/quadratures.py:56-95)
# --------------------------------------------------------------------------------------
@dataclass
class Face:
    L: np.ndarray
    theta: np.ndarray
    nx: np.ndarray
    ny: np.ndarray
    mx: np.ndarray   # midpoint x
    my: np.ndarray   # midpoint y
    qx: np.ndarray   # 1-point Gauss point x
    qy: np.ndarray


# Gauss rules on [-1, 1] as (point, weight) in the reference's order and with its expressions
# (mesh/quadratures.py:33-37): the weights sum to 2; integrate_flux multiplies by L and _dUdt_JIT by 0.5.
QUADS = {
    1: ((0, 2),),
    2: ((-1 / np.sqrt(3), 1), (1 / np.sqrt(3), 1)),
    3: ((-np.sqrt(3 / 5), 5 / 9), (0, 8 / 9), (np.sqrt(3 / 5), 5 / 9)),
}


@dataclass
class Geometry:
    nx: int
    ny: int
    xn: np.ndarray        # node x (ny+1, nx+1)
    yn: np.ndarray
    xc: np.ndarray        # centroid (ny, nx)
    yc: np.ndarray
    A: np.ndarray
    dx: np.ndarray
    dy: np.ndarray
    face: Dict[str, Face]
    is_cartesian: bool
    vertices: dict
    # Gauss points per face: [(x, y, w)] in rule order (QuadraturePointData, mesh/quadratures.py:60-95)
    qpts: Dict[str, list] = None


def _cellface(hx, hy, lx, ly, vertical: bool, direction: int):
    """CellFace (mesh/base.py:29-78): length, midpoint, angle and normal."""
    L = np.sqrt((hx - lx) ** 2 + (hy - ly) ** 2)
    mx = 0.5 * (hx + lx)
    my = 0.5 * (hy + ly)
    with np.errstate(divide="ignore", invalid="ignore"):
        if vertical:
            theta = np.arctan((lx - hx) / (hy - ly))
        else:
            theta = np.pi / 2 - np.arctan((hy - ly) / (lx - hx))
    nx = direction * np.cos(theta)
    ny = direction * np.sin(theta)
    return L, theta, nx, ny, mx, my


def _face_length_jit(x1, x2, y1, y2):
    """QuadMesh._face_length (mesh/quad_mesh.py:258-268)."""
    dx = x1 - x2
    dy = y1 - y2
    return np.sqrt(dx * dx + dy * dy)


def _qp(p1, p2, p=0):
    """QuadraturePointData._transform (mesh/quadratures.py:56-58) at the rule point p."""
    return 0.5 * ((p2 - p1) * p + (p2 + p1))


def block_geometry(vert: dict, nx: int, ny: int, nq: int = 1) -> Geometry:
    """QuadMesh.create_mesh (mesh/quad_mesh.py:71-135) and the nq-point Gauss points of every face
    (QuadraturePointData._make_quadrature, mesh/quadratures.py:60-95)."""
    SW, SE, NW, NE = vert["SW"], vert["SE"], vert["NW"], vert["NE"]
    Ex = np.linspace(SE[0], NE[0], ny + 1)
    Ey = np.linspace(SE[1], NE[1], ny + 1)
    Wx = np.linspace(SW[0], NW[0], ny + 1)
    Wy = np.linspace(SW[1], NW[1], ny + 1)
    x = np.zeros((ny + 1, nx + 1))
    y = np.zeros((ny + 1, nx + 1))
    for i in range(ny + 1):
        x[i, :] = np.linspace(Wx[i], Ex[i], nx + 1)
        y[i, :] = np.linspace(Wy[i], Ey[i], nx + 1)
    xne, yne = x[1:, 1:], y[1:, 1:]
    xnw, ynw = x[1:, :-1], y[1:, :-1]
    xse, yse = x[:-1, 1:], y[:-1, 1:]
    xsw, ysw = x[:-1, :-1], y[:-1, :-1]
    # centroid (quad_mesh.py:186-199)
    xc = 0.25 * (xne + xnw + xse + xsw)
    yc = 0.25 * (yne + ynw + yse + ysw)
    faces = {}
    # (high, low) node pairs (quad_mesh.py:105-128)
    spec = {
        "E": (xne, yne, xse, yse, True, 1),
        "W": (xnw, ynw, xsw, ysw, True, -1),
        "N": (xnw, ynw, xne, yne, False, 1),
        "S": (xsw, ysw, xse, yse, False, -1),
    }
    qps = {
        "E": (_qp(xne, xse), _qp(yne, yse)),
        "W": (_qp(xnw, xsw), _qp(ynw, ysw)),
        "N": (_qp(xne, xnw), _qp(yne, ynw)),
        "S": (_qp(xse, xsw), _qp(yse, ysw)),
    }
    for d, (hx, hy, lx, ly, vert_, dr) in spec.items():
        L, th, fnx, fny, mx, my = _cellface(hx, hy, lx, ly, vert_, dr)
        faces[d] = Face(L, th, fnx, fny, mx, my, qps[d][0], qps[d][1])
    # area (quad_mesh.py:137-184)
    s1 = _face_length_jit(xsw, xnw, ynw, ysw)
    s3 = _face_length_jit(xse, xne, yne, yse)
    s2 = _face_length_jit(xnw, xne, yne, ynw)
    s4 = _face_length_jit(xsw, xse, yse, ysw)
    d2 = (x[1:, :-1] - x[:-1, 1:]) ** 2 + (y[1:, :-1] - y[:-1, 1:]) ** 2
    with np.errstate(invalid="ignore"):
        a1 = np.arccos((s1 ** 2 + s4 ** 2 - d2) / (2 * s1 * s4))
        a2 = np.arccos((s2 ** 2 + s3 ** 2 - d2) / (2 * s2 * s3))
    s = 0.5 * (s1 + s2 + s3 + s4)
    p1 = (s - s1) * (s - s2) * (s - s3) * (s - s4)
    p2 = s1 * s2 * s3 * s4
    A = np.sqrt(p1 - 0.5 * p2 * (1 + np.cos(a1 + a2)))
    dx = faces["E"].mx - faces["W"].mx
    dy = faces["N"].my - faces["S"].my
    cart = (NE[1] == NW[1]) and (SE[1] == SW[1]) and (SE[0] == NE[0]) and (SW[0] == NW[0])
    # (p1, p2) vertex pairs of each face: E (NE, SE), W (NW, SW), N (NE, NW), S (SE, SW)
    ends = {"E": ((xne, yne), (xse, yse)), "W": ((xnw, ynw), (xsw, ysw)),
            "N": ((xne, yne), (xnw, ynw)), "S": ((xse, yse), (xsw, ysw))}
    qpts = {d: [(_qp(a[0], b[0], p), _qp(a[1], b[1], p), w) for p, w in QUADS[nq]]
            for d, (a, b) in ends.items()}
    return Geometry(nx, ny, x, y, xc, yc, A, dx, dy, faces, bool(cart), vert, qpts)


# --------------------------------------------------------------------------------------
# Rotations (utils/utils.py:75-85,88-112,148-158,161-185) -- libm sin/cos like Numba
# --------------------------------------------------------------------------------------
_vsin = np.vectorize(math.sin, otypes=[float])
_vcos = np.vectorize(math.cos, otypes=[float])


def rotate(theta, arr):
    """rotate_JIT: in place, theta broadcastable to arr[..., 0]."""
    th = np.broadcast_to(theta, arr.shape[:2])
    s = _vsin(th)
    c = _vcos(th)
    u = arr[:, :, 1].copy()
    v = arr[:, :, 2].copy()
    arr[:, :, 1] = u * c + v * s
    arr[:, :, 2] = v * c - u * s


def unrotate(theta, arr):
    """unrotate_JIT: in place."""
    th = np.broadcast_to(theta, arr.shape[:2])
    s = _vsin(th)
    c = _vcos(th)
    u = arr[:, :, 1].copy()
    v = arr[:, :, 2].copy()
    arr[:, :, 1] = u * c - v * s
    arr[:, :, 2] = v * c + u * s


def rotate90(*arrays):
    for a in arrays:
        u = a[:, :, 2].copy()
        v = a[:, :, 1].copy()
        a[:, :, 1], a[:, :, 2] = u, -v


def unrotate90(*arrays):
    for a in arrays:
        u = a[:, :, 2].copy()
        v = a[:, :, 1].copy()
        a[:, :, 1], a[:, :, 2] = -u, v


def reflect(state, wall_theta):
    """BoundaryConditionFunctions.reflection (boundary_conditions/funcs.py:27-43)."""
    rotate(wall_theta, state)
    state[:, :, 1] = -state[:, :, 1]
    unrotate(wall_theta, state)


# --------------------------------------------------------------------------------------
# State conversions (states/conservative.py:85-137, states/primitive.py:93-237,
# states/converter/concrete_defs.py:85-141)
# --------------------------------------------------------------------------------------
def cons_to_prim(U):
    rho = U[:, :, 0]
    u = U[:, :, 1] / rho
    v = U[:, :, 2] / rho
    ek = rho * (0.5 * (u * u + v * v))
    p = GM1 * (U[:, :, 3] - ek)
    return np.dstack((rho.copy(), u, v, p))


def prim_ek(W):
    """PrimitiveState.ek_JIT (states/primitive.py:93-104)."""
    return 0.5 * W[:, :, 0] * (W[:, :, 1] * W[:, :, 1] + W[:, :, 2] * W[:, :, 2])


def prim_to_cons(W):
    rho, u, v, p = W[:, :, 0], W[:, :, 1], W[:, :, 2], W[:, :, 3]
    return np.dstack((rho.copy(), rho * u, rho * v, p / GM1 + prim_ek(W)))


def prim_a(W):
    """PrimitiveState.a_JIT (states/primitive.py:178-189)."""
    return np.sqrt(GAMMA * W[:, :, 3] / W[:, :, 0])


def prim_F(W):
    """PrimitiveState._F_from_prim_JIT (states/primitive.py:223-237)."""
    ek = prim_ek(W)
    u = W[:, :, 1]
    ru = W[:, :, 0] * u
    F = np.zeros(W.shape)
    F[:, :, 0] = ru
    F[:, :, 1] = ru * u + W[:, :, 3]
    F[:, :, 2] = ru * W[:, :, 2]
    F[:, :, 3] = u * (ONE_OVER_GM1 * W[:, :, 3] + ek + W[:, :, 3])
    return F


def roe_state(WL, WR):
    """RoePrimitiveState._roe_state_from_prim_JIT (states/primitive.py:301-320)."""
    sL = np.sqrt(WL[:, :, 0])
    sR = np.sqrt(WR[:, :, 0])
    inv = 1.0 / (sL + sR)
    Q = np.zeros(WL.shape)
    Q[:, :, 0] = np.sqrt(WL[:, :, 0] * WR[:, :, 0])
    for k in (1, 2, 3):
        Q[:, :, k] = (WL[:, :, k] * sL + WR[:, :, k] * sR) * inv
    return Q


def harten(R_p, R_m, L_p, L_m, Roe_p, Roe_m):
    """FluxFunction._harten_correction_JIT (flux/base.py:119-146); returns corrected copies."""
    tp = 2 * (R_p - L_p)
    tm = 2 * (R_m - L_m)
    tp = np.where(tp <= 0, 1e-8, tp)
    tm = np.where(tm <= 0, 1e-8, tm)
    rp = np.where(np.absolute(Roe_p) < tp, 0.5 * ((Roe_p * Roe_p) / tp + tp), Roe_p)
    rm = np.where(np.absolute(Roe_m) < tm, 0.5 * ((Roe_m * Roe_m) / tm + tm), Roe_m)
    return rp, rm


def wavespeeds(W, a=None):
    """FluxFunction._wavespeeds_x_JIT (flux/base.py:64-75): (u - a, u + a)."""
    if a is None:
        a = prim_a(W)
    return W[:, :, 1] - a, W[:, :, 1] + a


# ---- Roe eigen-system COO layout (flux/eigen_system.py:62-252) ----
def _blk_idx(F, rows):
    return np.concatenate([np.arange(n, 4 * F, 4) for n in rows])


def _roe_coo(F):
    """Index arrays of Rc, Lp and Lambda in the reference's storage order."""
    Rc_i = np.hstack([_blk_idx(F, r) for r in ([3], [2, 3], [1, 2, 3], [0, 1, 2, 3], [0, 2], [1], [0])])
    Rc_j = np.hstack([_blk_idx(F, c) for c in ([0], [0, 1], [0, 1, 2], [0, 1, 2, 3], [1, 3], [3], [3])])
    Lp_i = np.hstack([_blk_idx(F, r) for r in ([3], [1], [2, 3], [0], [1], [0])])
    Lp_j = np.hstack([_blk_idx(F, c) for c in ([1], [0], [2, 3], [1], [3], [3])])
    L_i = _blk_idx(F, [0, 1, 2, 3])
    return Rc_i, Rc_j, Lp_i, Lp_j, L_i


def roe_rc_data(Wr, aroe, Lm, Lp):
    """FluxRoe.compute_Rc (flux/Roe.py:175-220): the 14 COO blocks of Rc, with the Harten-corrected
    slow/fast speeds Lm, Lp in the (1,0) and (1,3) entries (quirk B.2); unset blocks keep the 1.0 of
    np.ones (flux/eigen_system.py:36-42)."""
    num = Wr.shape[0] * Wr.shape[1]
    u = Wr[:, :, 1]
    v = Wr[:, :, 2]
    Ek = 0.5 * (u * u + v * v)                       # Ek_JIT
    H = G_OVER_GM1 * Wr[:, :, 3] / Wr[:, :, 0] + Ek  # H_given_Ek_JIT
    ua = u * aroe
    Rc_data = np.ones(14 * num)
    seg = lambda k, val: Rc_data.__setitem__(slice(k * num, (k + 1) * num), np.ravel(val))
    seg(0, H - ua); seg(1, v); seg(2, Ek); seg(3, Lm); seg(4, v); seg(5, v)
    seg(7, u); seg(9, H + ua); seg(11, v); seg(12, Lp)
    return Rc_data


def roe_lp_data(Wr, aroe):
    """FluxRoe.compute_Lp (flux/Roe.py:222-251): the 7 COO blocks of Lp.  Block 2, the (2,2) entry,
    is never written and keeps the 1.0 of np.ones (flux/eigen_system.py:52-57): the shear-wave
    strength is dv, not rho* dv (quirk B.15)."""
    num = Wr.shape[0] * Wr.shape[1]
    _a = 1.0 / aroe
    one_a2 = _a * _a
    h_one_a2 = 0.5 * one_a2
    r2a = 0.5 * Wr[:, :, 0] * _a
    Lp_data = np.ones(7 * num)
    for k, val in ((0, r2a), (3, h_one_a2), (4, -r2a), (5, -one_a2), (6, h_one_a2)):
        Lp_data[k * num:(k + 1) * num] = np.ravel(val)
    return Lp_data


def flux_roe(WL, WR):
    """FluxRoe.compute_flux (flux/Roe.py:262-304) on primitive (m, F, 4) face arrays: central flux
    minus 0.5 Rc |Lambda| Lp dW with the Roe state of states/primitive.py:301-320, Harten-corrected
    eigenvalues (flux/Roe.py:292-304, flux/base.py:119-146) and SciPy COO mat-vecs in the
    reference's storage order."""
    shape = WL.shape
    F = shape[0] * shape[1]
    wl = WL.reshape(1, F, 4)
    wr = WR.reshape(1, F, 4)
    Wr = roe_state(wl, wr)
    aroe = prim_a(Wr)
    R_p, R_m = wavespeeds(wr)
    L_p, L_m = wavespeeds(wl)
    Roe_p, Roe_m = wavespeeds(Wr, aroe)
    Lm, Lp = harten(R_p, R_m, L_p, L_m, Roe_p, Roe_m)   # slow, fast (corrected)
    u = Wr[:, :, 1]
    Rc_i, Rc_j, Lp_i, Lp_j, L_i = _roe_coo(F)
    Rc_data = roe_rc_data(Wr, aroe, Lm, Lp)
    Lp_data = roe_lp_data(Wr, aroe)
    L_data = np.concatenate([np.ravel(Lm), np.ravel(u), np.ravel(u), np.ravel(Lp)])
    Rc = coo_matrix((Rc_data, (Rc_i, Rc_j)))
    Lpm = coo_matrix((Lp_data, (Lp_i, Lp_j)))
    Lam = coo_matrix((np.absolute(L_data), (L_i, L_i)))
    dW = (wr - wl).flatten()
    up = 0.5 * (Rc * (Lam * (Lpm * dW))).reshape(1, -1, 4)
    flux = 0.5 * (prim_F(wl) + prim_F(wr)) - up
    return flux.reshape(shape)


def nrm2(x):
    """Euclidean norm over the last axis as flux/HLLL.py:86-93 computes it: Numba's np.linalg.norm
    of a 1-D float64 array calls the BLAS dnrm2 of its SciPy (OpenBLAS 0.3.28 with scipy 1.15.3),
    whose x86-64 kernel sums 0 + x0^2 + x1^2 + x2^2 + x3^2 in order in the x87 80-bit format, takes
    the square root there and rounds once to double.  np.longdouble is that format on x86-64 Linux;
    pinned against scipy.linalg.blas.dnrm2 in tests/test_oracle_pins.py."""
    if np.finfo(np.longdouble).nmant != 63:
        raise RuntimeError("nrm2 needs the x87 80-bit long double (x86-64)")
    X = np.asarray(x, dtype=np.float64).astype(np.longdouble)
    s = np.zeros(X.shape[:-1], dtype=np.longdouble)
    for k in range(X.shape[-1]):
        s = s + X[..., k] * X[..., k]
    return np.sqrt(s).astype(np.float64)


def hlll_coef(aroe, u, dU, dF, Lminus, Lplus):
    """_HLLL_flux_JIT's dissipation coefficient (flux/HLLL.py:84-101): alpha = max(0, 1 - |dF - u dU|
    / k) with k = a |dU| (k + 1e-14 when k < 1e-16), then Lm Lp (1 - alpha (1 - max(u/Lm, u/Lp)))."""
    k = aroe * nrm2(dU)
    n2 = nrm2(dF - u[:, :, None] * dU)
    alpha = np.maximum(0, 1 - n2 / np.where(k < 1e-16, k + 1e-14, k))
    mx = np.maximum(u / Lminus, u / Lplus)
    return Lminus * Lplus * (1 - alpha * (1 - mx))


def flux_hlll(WL, WR):
    """FluxHLLL.compute_flux + _HLLL_flux_JIT (flux/HLLL.py:25-49,70-103)."""
    Wr = roe_state(WL, WR)
    aroe = prim_a(Wr)
    L_p, L_m = wavespeeds(WL)
    R_p, R_m = wavespeeds(WR)
    Roe_p, Roe_m = wavespeeds(Wr, aroe)
    c_slow, c_fast = harten(R_p, R_m, L_p, L_m, Roe_p, Roe_m)
    Lplus = np.maximum(R_m, c_fast)
    Lminus = np.minimum(L_p, c_slow)
    UR = prim_to_cons(WR)
    UL = prim_to_cons(WL)

    def Fcons(W, U):  # PrimitiveState.F(U=...) (states/primitive.py:203-209)
        F = np.zeros(W.shape)
        F[:, :, 0] = U[:, :, 1]
        F[:, :, 1] = U[:, :, 1] * W[:, :, 1] + W[:, :, 3]
        F[:, :, 2] = U[:, :, 1] * W[:, :, 2]
        F[:, :, 3] = W[:, :, 1] * (U[:, :, 3] + W[:, :, 3])
        return F

    FR = Fcons(WR, UR)
    FL = Fcons(WL, UL)
    u = Wr[:, :, 1]
    dU = UR - UL
    dF = FR - FL
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        coef = hlll_coef(aroe, u, dU, dF, Lminus, Lplus)
        mid = (Lplus[:, :, None] * FL - Lminus[:, :, None] * FR + coef[:, :, None] * dU) / (Lplus - Lminus)[:, :, None]
    out = np.where((Lminus >= 0)[:, :, None], FL, np.where((Lplus <= 0)[:, :, None], FR, mid))
    return out


# --------------------------------------------------------------------------------------
# Blocks and the residual
# --------------------------------------------------------------------------------------
@dataclass
class Block:
    num: int
    geom: Geometry
    neighbors: Dict[str, Optional[int]]
    bc: Dict[str, Optional[str]]
    U: np.ndarray = None
    ghost: Dict[str, np.ndarray] = field(default_factory=dict)
    limiter_fn: object = None          # limiter function (None: Venkatakrishnan)
    prim: bool = False                 # reconstruction_type is PrimitiveState


def boundary_theta(g: Geometry, d: str):
    """QuadMesh.boundary_angle (mesh/quad_mesh.py:362-369)."""
    th = g.face[d].theta
    return {"E": th[:, -1:], "W": th[:, :1], "N": th[-1:, :], "S": th[:1, :]}[d]


_ghost_sl = {"E": np.s_[:, -1:, :], "W": np.s_[:, :1, :], "N": np.s_[-1:, :, :], "S": np.s_[:1, :, :]}


def apply_boundary_condition(blocks: Dict[int, Block]):
    """Blocks.apply_boundary_condition (blocks/base.py:448-471) with
    GhostBlock.send_boundary_data (blocks/ghost.py:187-216) and the BC functions (:243-278)."""
    for b in blocks.values():
        for d in DIRS:
            nb = b.neighbors[d]
            if b.bc[d] is not None or nb is None:
                b.ghost[d] = b.U[_ghost_sl[d]].copy()
            else:
                b.ghost[d] = blocks[nb].U[_ghost_sl[OPPOSITE[d]]].copy()
    for b in blocks.values():
        for d in DIRS:
            if b.bc[d] in ("Reflection", "Slipwall"):
                reflect(b.ghost[d], boundary_theta(b.geom, d))
            elif b.bc[d] not in (None, "OutletDirichlet"):
                raise ValueError(f"Boundary Condition type {b.bc[d]} has not been specialized.")


def interface_values(U, gE, gW, gN, gS):
    """_get_interface_values_arithmetic_JIT (blocks/quad_block.py:181-218)."""
    ny, nx, _ = U.shape
    IEW = np.zeros((ny, nx + 1, 4))
    INS = np.zeros((ny + 1, nx, 4))
    IEW[:, 0] = 0.5 * (gW[:, 0] + U[:, 0])
    IEW[:, -1] = 0.5 * (gE[:, 0] + U[:, -1])
    INS[0] = 0.5 * (gS[0] + U[0])
    INS[-1] = 0.5 * (gN[0] + U[-1])
    IEW[:, 1:nx] = 0.5 * (U[:, :-1] + U[:, 1:])
    INS[1:ny] = 0.5 * (U[:-1] + U[1:])
    return IEW[:, 1:], IEW[:, :-1], INS[1:], INS[:-1]


def green_gauss(U, ghost, g: Geometry):
    """GreenGauss._get_gradinet_JIT (gradients/greengauss.py:110-155)."""
    e, w, n, s = interface_values(U, ghost["E"], ghost["W"], ghost["N"], ghost["S"])
    f = g.face
    xl = {d: f[d].L * f[d].nx for d in DIRS}
    yl = {d: f[d].L * f[d].ny for d in DIRS}
    a = 1.0 / g.A
    gx = (e * xl["E"][..., None] + w * xl["W"][..., None] + n * xl["N"][..., None] + s * xl["S"][..., None]) * a[..., None]
    gy = (e * yl["E"][..., None] + w * yl["W"][..., None] + n * yl["N"][..., None] + s * yl["S"][..., None]) * a[..., None]
    return gx, gy


def high_order_term(gx, gy, g: Geometry, d: str, i: int = 0):
    """FirstOrderGradients.get_high_order_term_JIT (blocks/base.py:261-289) at Gauss point i of
    face d."""
    qx, qy, _ = g.qpts[d][i]
    ddx = qx - g.xc
    ddy = qy - g.yc
    return gx * ddx[..., None] + gy * ddy[..., None]


def venkatakrishnan(s):
    """Venkatakrishnan._venkata (limiters/limiters.py:38-49)."""
    s2 = s * s
    return (s2 + 2 * s) / (s2 + s + 2)


def barth_jespersen(s):
    """BarthJespersen._limiter_func (limiters/limiters.py:25-28): np.minimum(1, s)."""
    return np.minimum(1, s)


def van_albada(s):
    """VanAlbada._limiter_func (limiters/limiters.py:52-56); slope**2 is s*s."""
    s2 = s * s
    return (s2 + s) / (s2 + 1)


def van_leer(s):
    """VanLeer._limiter_func (limiters/limiters.py:59-62)."""
    return (np.absolute(s) + s) / (s + 1)


# SlopeLimiterFactory.create (limiters/factory.py:28-39)
LIMITERS = {"Venkatakrishnan": venkatakrishnan, "BarthJespersen": barth_jespersen, "VanAlbada": van_albada,
            "VanLeer": van_leer}


def compute_slope(dmax, dmin, d):
    """SlopeLimiter._compute_slope (limiters/base.py:189-221)."""
    s = np.ones_like(d)
    with np.errstate(divide="ignore", invalid="ignore"):
        pos = dmax / d
        neg = dmin / d
    s = np.where(d > 0, pos, s)
    s = np.where(d < 0, neg, s)
    return s


def reconstruction_delta(U, h):
    """SecondOrderMUSCL.unlimited_solution_at_quadrature_point minus the cell average
    (fvm/SecondOrderMUSCL.py:86-94, limiters/base.py:96-107): (U + hot) - U, which is 0, not hot,
    wherever hot is below half an ulp of U (quirk B.4)."""
    q = U + h           # unlimited solution at the quadrature point
    return q - U


def limiter_phi(U, ghost, hot: Dict[str, object], limiter_fn=venkatakrishnan):
    """SlopeLimiter._get_slope / _limit (limiters/base.py:47-108,147-187) with the limiter
    function of the configured limiter (default Venkatakrishnan).  hot[d]: the high-order term of
    face d, or a list of them, one per Gauss point; phi is the minimum over the faces' points in
    the order E, W, N, S (each face's points in rule order)."""
    EW = np.concatenate((ghost["W"], U, ghost["E"]), axis=1)
    NS = np.concatenate((ghost["S"], U, ghost["N"]), axis=0)
    vals = (U, EW[:, :-2], EW[:, 2:], NS[:-2, :], NS[2:, :])
    dmax = np.maximum.reduce(vals) - U
    dmin = np.minimum.reduce(vals) - U
    phis = []
    for d in DIRS:
        for h in (hot[d] if isinstance(hot[d], (list, tuple)) else (hot[d],)):
            dd = reconstruction_delta(U, h)
            with np.errstate(divide="ignore", invalid="ignore"):
                phis.append(limiter_fn(compute_slope(dmax, dmin, dd)))
    phi = np.minimum.reduce(tuple(phis))
    return np.where(phi < 0, 0, phi)


def _outer_state(b: Block, d: str, limited_boundary: np.ndarray, ghost=None):
    """MUSCL._get_boundary_flux_states (fvm/base.py:305-325) + the BC applied to it
    (fvm/base.py:352-362,421-431).  `ghost`: the reconstruction block's ghost states (default: the
    block's conservative ghosts)."""
    ghost = b.ghost if ghost is None else ghost
    if b.bc[d] is None:
        # ghost block fvm: phi = 0, gradient = 0 -> ghost cell average (+0)
        return ghost[d].copy() + 0.0
    st = limited_boundary.copy()
    if b.bc[d] in ("Reflection", "Slipwall"):
        reflect(st, boundary_theta(b.geom, d))
    return st


def residual(b: Block, flux_fn):
    """FiniteVolumeMethod.dUdt (fvm/base.py:108-165) for SecondOrderMUSCL (MUSCL.evaluate_flux,
    fvm/base.py:327-500) with the block geometry's Gauss rule: one flux evaluation per face and
    Gauss point, the points of a face paired in rule order (fvm/base.py:349-356,411-418)."""
    # recon_block.from_block (blocks/quad_block.py:120-134): the interior and ghost states in the
    # reconstruction type -- a copy (ConservativeState) or ConservativeConverter.to_primitive
    # (PrimitiveState; concrete_defs.py:85-100)
    if b.prim:
        U = cons_to_prim(b.U)
        ghost = {d: cons_to_prim(b.ghost[d]) for d in DIRS}
    else:
        U = b.U.copy()
        ghost = b.ghost
    # MUSCL._get_left_right_riemann_states (fvm/base.py:255-303): primitive face states go to the
    # flux function as they are
    to_flux = (lambda a: a) if b.prim else cons_to_prim
    g = b.geom
    nq = len(g.qpts["E"])
    gx, gy = green_gauss(U, ghost, g)
    hot = {d: [high_order_term(gx, gy, g, d, i) for i in range(nq)] for d in DIRS}
    phi = limiter_phi(U, ghost, hot, b.limiter_fn or venkatakrishnan)
    fluxes = {d: [] for d in DIRS}
    for i in range(nq):
        # limited_solution_at_quadrature_point at Gauss point i of every face
        face = {d: U + phi * hot[d][i] for d in DIRS}
        bnd = {"E": face["E"][:, -1:], "W": face["W"][:, :1], "N": face["N"][-1:], "S": face["S"][:1]}
        # ---- east-west sweep (fvm/base.py:327-390)
        outE = _outer_state(b, "E", bnd["E"], ghost)
        outW = _outer_state(b, "W", bnd["W"], ghost)
        fE = face["E"].copy()
        fW = face["W"].copy()
        if not g.is_cartesian:
            rotate(g.face["E"].theta, fE)
            rotate(g.face["W"].theta, fW)
            rotate(boundary_theta(g, "E"), outE)
            rotate(boundary_theta(g, "W"), outW)
        left = to_flux(np.concatenate((outW, fE), axis=1))
        right = to_flux(np.concatenate((fW, outE), axis=1))
        Fx = flux_fn(left, right)
        flux_E = Fx[:, 1:].copy()
        flux_W = Fx[:, :-1].copy()
        if not g.is_cartesian:
            unrotate(g.face["E"].theta, flux_E)
            unrotate(g.face["W"].theta, flux_W)
        # ---- north-south sweep (fvm/base.py:392-486)
        outN = _outer_state(b, "N", bnd["N"], ghost)
        outS = _outer_state(b, "S", bnd["S"], ghost)
        fN = face["N"].copy()
        fS = face["S"].copy()
        if g.is_cartesian:
            rotate90(fN, fS, outN, outS)
        else:
            rotate(g.face["N"].theta, fN)
            rotate(g.face["S"].theta, fS)
            rotate(boundary_theta(g, "N"), outN)
            rotate(boundary_theta(g, "S"), outS)
        left = to_flux(np.concatenate((outS.transpose(1, 0, 2), fN.transpose(1, 0, 2)), axis=1))
        right = to_flux(np.concatenate((fS.transpose(1, 0, 2), outN.transpose(1, 0, 2)), axis=1))
        Fy = flux_fn(left, right).transpose(1, 0, 2)
        flux_N = Fy[1:].copy()
        flux_S = Fy[:-1].copy()
        if g.is_cartesian:
            unrotate90(flux_N, flux_S)
        else:
            unrotate(g.face["N"].theta, flux_N)
            unrotate(g.face["S"].theta, flux_S)
        for d, fl in (("E", flux_E), ("W", flux_W), ("N", flux_N), ("S", flux_S)):
            fluxes[d].append(fl)
    # ---- integrate_flux: L * sum(w_i * F_i), the sum starting at 0 (fvm/base.py:168-190);
    # dUdt (fvm/base.py:118-139,141-165)
    I = {}
    for d in DIRS:
        acc = 0
        for (_, _, w), fl in zip(g.qpts[d], fluxes[d]):
            acc = acc + w * fl
        I[d] = g.face[d].L[..., None] * acc
    return 0.5 * (I["W"] - I["E"] + I["S"] - I["N"]) / g.A[..., None]


# --------------------------------------------------------------------------------------
# Time marching, dt, realizability
# --------------------------------------------------------------------------------------
BUTCHER = {
    # time_marching/explicit_runge_kutta.py:91-274
    "ExplicitEuler1": [[1]],
    "RK2": [[0.5], [0, 1]],
    "Generic2": None,           # needs alpha; SSPRK2 = Generic2(alpha=1)
    "SSPRK2": [[1], [1 - 1 / (2 * 1), 1 / (2 * 1)]],
    "Ralston2": [[2 / 3], [1 / 4, 3 / 4]],
    "RK3": [[0.5], [-1, 2], [1 / 6, 2 / 3, 1 / 6]],
    "RK3SSP": [[1], [1 / 4, 1 / 4], [1 / 6, 1 / 6, 2 / 3]],
    "Ralston3": [[1 / 2], [0, 3 / 4], [2 / 9, 1 / 3, 4 / 9]],
    "RK4": [[0.5], [0, 0.5], [0, 0, 1], [1 / 6, 1 / 3, 1 / 3, 1 / 6]],
    "Ralston4": [[0.4], [0.29697761, 0.15875964], [0.21810040, -3.05096516, 3.83286476],
                 [0.17476028, -0.55148066, 1.20553560, 0.17118478]],
    "DormandPrince5": [[1 / 5], [3 / 40, 9 / 40], [44 / 45, -56 / 15, 32 / 9],
                       [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
                       [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
                       [35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84]],
}


def generic2(a):
    """ExplicitRungeKutta.Generic2 (time_marching/explicit_runge_kutta.py:103-115)."""
    return [[a], [1 - 1 / (2 * a), 1 / (2 * a)]]


def generic3(a):
    """ExplicitRungeKutta.Generic3 (time_marching/explicit_runge_kutta.py:143-168)."""
    if a in (0, 2 / 3, 1):
        raise ValueError("Value of alpha parameter is not allowd.")
    k = (1 - a) / a / (3 * a - 2)
    return [[a], [1 + k, -k], [0.5 - 1 / (6 * a), 1 / (6 * a * (1 - a)), (2 - 3 * a) / (6 * (1 - a))]]


def block_dt(b: Block, CFL: float) -> float:
    """QuadBlock.get_dt (blocks/quad_block.py:423-436) on the conservative state."""
    U = b.U
    rho = U[:, :, 0]
    u = U[:, :, 1] / rho
    v = U[:, :, 2] / rho
    p = GM1 * (U[:, :, 3] - rho * (0.5 * (u * u + v * v)))
    a = np.sqrt(GAMMA * p / rho)
    tx = b.geom.dx / (np.absolute(u) + a)
    ty = b.geom.dy / (np.absolute(v) + a)
    return CFL * np.amin(np.minimum(tx, ty))


def realizable(blocks: Dict[int, Block]) -> bool:
    """Blocks.realizable (blocks/base.py:442-446; states/conservative.py:161-165)."""
    for b in blocks.values():
        arrs = [b.U] + [b.ghost[d] for d in DIRS]
        for a in arrs:
            if not (np.all(a[:, :, 0] > 0) and np.all(a[:, :, 3] > 0)):
                return False
    return True


def integrate(blocks: Dict[int, Block], dt: float, a: List[List[float]], flux_fn):
    """ExplicitRungeKutta.integrate (time_marching/explicit_runge_kutta.py:47-89)."""
    k = {i: {} for i in blocks}
    U0 = {i: b.U.copy() for i, b in blocks.items()}
    temp = {i: np.zeros_like(b.U) for i, b in blocks.items()}
    for stage in range(len(a)):
        for i, b in reversed(list(blocks.items())):
            k[i][stage] = residual(b, flux_fn)
            inter = U0[i]
            for step in range(stage + 1):
                if a[stage][step] != 0:
                    inter = _axpy(temp[i], inter, dt * a[stage][step], k[i][step])
            b.U = inter
        apply_boundary_condition(blocks)


def _axpy(out, x, c, r):
    """ExplicitRungeKutta._update_state (explicit_runge_kutta.py:82-89)."""
    out[...] = x + c * r
    return out


class Simulation:
    """Step-count loop with the semantics of Euler2D._solve (solvers/Euler2D.py:175-216)
    and Solver.get_dt (solvers/base.py:114-136)."""

    def __init__(self, mesh: Dict[int, dict], nx: int, ny: int, *, flux: str = "Roe",
                 integrator="RK4", CFL: float = 0.8, t_final: float = 1e30, a_inf: float = 1.0,
                 limiter: str = "Venkatakrishnan", reconstruction: str = "Conservative",
                 quadrature_points: int = 1):
        self.flux_fn = {"Roe": flux_roe, "HLLL": flux_hlll}[flux]
        self.limiter_fn = LIMITERS[limiter]
        self.a = BUTCHER[integrator] if isinstance(integrator, str) else integrator
        self.CFL = CFL
        self.t = 0.0
        self.t_final = t_final * a_inf
        self.nsteps = 0
        self.dts: List[float] = []
        self.blocks: Dict[int, Block] = {}
        for n, d in mesh.items():
            vert = {k: d[k] for k in ("SW", "SE", "NW", "NE")}
            g = block_geometry(vert, nx, ny, quadrature_points)
            self.blocks[n] = Block(n, g, {s: d["Neighbor" + s] for s in DIRS},
                                   {s: d["BCType" + s] for s in DIRS}, limiter_fn=self.limiter_fn,
                                   prim=(reconstruction == "Primitive"))

    def set_state(self, states: Dict[int, np.ndarray]):
        for n, b in self.blocks.items():
            b.U = np.array(states[n], dtype=np.float64, copy=True)
        apply_boundary_condition(self.blocks)

    def get_dt(self) -> float:
        dt = min(block_dt(b, self.CFL) for b in self.blocks.values())
        return self.t_final - self.t if self.t_final - self.t < dt else dt

    def step(self, dt: Optional[float] = None):
        if dt is None:
            dt = self.get_dt()
        integrate(self.blocks, dt, self.a, self.flux_fn)
        if not realizable(self.blocks):
            raise FloatingPointError(f"unrealizable state after step {self.nsteps + 1}")
        self.dts.append(dt)
        self.t += dt
        self.nsteps += 1
        return dt

    def run(self, nsteps: int, dts: Optional[List[float]] = None):
        for s in range(nsteps):
            self.step(None if dts is None else dts[s])

    def states(self) -> Dict[int, np.ndarray]:
        return {n: b.U.copy() for n, b in self.blocks.items()}


# --------------------------------------------------------------------------------------
# Initial conditions used by the benchmark configurations
# --------------------------------------------------------------------------------------
def explosion_state(geom: Geometry, a_inf=343.0, rho_inf=1.0, inside=(4.6968, 404400.0),
                    outside=(1.1742, 101100.0), box=(3.0, 7.0, 3.0, 7.0)):
    """ExplosionInitialCondition.apply_to_block (examples/explosion/initial_condition.py:32-61)
    followed by State.make_non_dimensional (states/base.py:93-97)."""
    def cons(rho, p):
        W = np.array([rho, 0.0, 0.0, p]).reshape((1, 1, 4))
        return prim_to_cons(W)
    left = cons(*inside)
    right = cons(*outside)
    x = geom.xc[..., None]
    y = geom.yc[..., None]
    xc = np.logical_and(x >= box[0], x <= box[1])
    yc = np.logical_and(y >= box[2], y <= box[3])
    U = np.where(np.logical_and(xc, yc), left, right)
    U[:, :, 0] /= rho_inf
    U[:, :, 1] /= rho_inf * a_inf
    U[:, :, 2] /= rho_inf * a_inf
    U[:, :, 3] /= rho_inf * a_inf ** 2
    return U
