"""Block meshes and their geometry, computed on the host.

* ``QuadMeshGenerator`` / ``RectagularMeshGenerator`` produce the reference's block dictionaries
  (mesh/base.py:133-226, mesh/rectangular.py:23-62), including the transfinite interpolation of
  interior block corners (mesh/base.py:88-121).
* ``BlockGeometry.build`` evaluates the per-cell geometry of one block with the reference's
  formulas and operation order (mesh/quad_mesh.py:71-212 nodes, centroids, Bretschneider area;
  mesh/base.py:29-78 face length, midpoint, angle, normal; mesh/quadratures.py:56-95 1-point Gauss
  points) and packs it into the layouts the kernels read (include/pyhype_hip.h).  Every sharing
  or separability the packed layout relies on is checked for exact equality against the per-cell
  arrays; a block that fails the Cartesian checks uses the general layout.
* Face-angle sines and cosines used by rotations are taken from libm (``math.sin``/``math.cos``),
  which is what the reference's Numba-compiled ``rotate_JIT`` calls; face normals use NumPy's
  ``np.cos``/``np.sin`` like ``CellFace._compute_face_norms``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

SIDES = ("E", "W", "N", "S")

# Gauss rules on [-1, 1] as (point, weight) pairs in the reference's order, with the reference's
# expressions (mesh/quadratures.py:33-37).
QUADS = {
    1: ((0, 2),),
    2: ((-1 / np.sqrt(3), 1), (1 / np.sqrt(3), 1)),
    3: ((-np.sqrt(3 / 5), 5 / 9), (0, 8 / 9), (np.sqrt(3 / 5), 5 / 9)),
}


class MeshGenerator:
    def __init__(self):
        self.dict = {}


class QuadMeshGenerator(MeshGenerator):
    """Block layout of an nx_blk x ny_blk patch (mesh/base.py:133-226)."""

    def __init__(self, nx_blk: int, ny_blk: int, BCE, BCW, BCN, BCS, BCNE=None, BCNW=None, BCSE=None,
                 BCSW=None, NE=None, NW=None, SE=None, SW=None, left_x=None, left_y=None, right_x=None,
                 right_y=None, top_x=None, top_y=None, bot_x=None, bot_y=None, blk_num_offset: int = 0):
        super().__init__()
        self.nx, self.ny = nx_blk + 1, ny_blk + 1
        self._blk_num_offset = blk_num_offset
        self.BCE = [BCE[0] for _ in range(ny_blk)] if len(BCE) == 1 else BCE
        self.BCW = [BCW[0] for _ in range(ny_blk)] if len(BCW) == 1 else BCW
        self.BCN = [BCN[0] for _ in range(nx_blk)] if len(BCN) == 1 else BCN
        self.BCS = [BCS[0] for _ in range(nx_blk)] if len(BCS) == 1 else BCS
        x, y = np.meshgrid(np.linspace(0, 1, self.nx), np.linspace(0, 1, self.ny))
        x[0, :] = np.linspace(SW[0], SE[0], self.nx) if bot_x is None else bot_x
        y[0, :] = np.linspace(SW[1], SE[1], self.nx) if bot_y is None else bot_y
        x[-1, :] = np.linspace(NW[0], NE[0], self.nx) if top_x is None else top_x
        y[-1, :] = np.linspace(NW[1], NE[1], self.nx) if top_y is None else top_y
        x[:, 0] = np.linspace(SW[0], NW[0], self.ny) if left_x is None else left_x
        y[:, 0] = np.linspace(SW[1], NW[1], self.ny) if left_y is None else left_y
        x[:, -1] = np.linspace(SE[0], NE[0], self.ny) if right_x is None else right_x
        y[:, -1] = np.linspace(SE[1], NE[1], self.ny) if right_y is None else right_y
        self.x, self.y = self._transfinite(x), self._transfinite(y)
        self.dict = self._create_block_descriptions()

    def _transfinite(self, a: np.ndarray) -> np.ndarray:
        n, m = self.nx, self.ny
        jm, im = np.meshgrid(np.linspace(1 / n, (n - 1) / n, n - 2), np.linspace(1 / m, (m - 1) / m, m - 2))
        mj, mi = np.meshgrid(np.linspace((n - 1) / n, 1 / n, n - 2), np.linspace((m - 1) / m, 1 / m, m - 2))
        a[1:-1, 1:-1] = (mi * a[0, 1:-1] + im * a[-1, 1:-1] + mj * a[1:-1, 0, None] + jm * a[1:-1, -1, None]
                         - mi * mj * a[0, 0] - mi * jm * a[0, -1] - im * mj * a[-1, 0] - im * jm * a[-1, -1])
        return a

    def _create_block_descriptions(self) -> Dict[int, dict]:
        ny, nx = self.ny - 1, self.nx - 1
        out = {}
        for i in range(ny):
            for j in range(nx):
                n = nx * i + j + self._blk_num_offset
                out[n] = {
                    "nBLK": n,
                    "NW": [self.x[i + 1, j], self.y[i + 1, j]],
                    "NE": [self.x[i + 1, j + 1], self.y[i + 1, j + 1]],
                    "SW": [self.x[i, j], self.y[i, j]],
                    "SE": [self.x[i, j + 1], self.y[i, j + 1]],
                    "NeighborE": n + 1 if j < nx - 1 else None,
                    "NeighborW": n - 1 if j > 0 else None,
                    "NeighborN": n + nx if n + nx < ny * nx else None,
                    "NeighborS": n - nx if n - nx >= 0 else None,
                    "NeighborNE": None, "NeighborNW": None, "NeighborSE": None, "NeighborSW": None,
                    "BCTypeE": self.BCE[i] if j == nx - 1 else None,
                    "BCTypeW": self.BCW[i] if j == 0 else None,
                    "BCTypeN": self.BCN[j] if i == ny - 1 else None,
                    "BCTypeS": self.BCS[j] if i == 0 else None,
                    "BCTypeNE": self.BCE[i] if j == nx - 1 else None,
                    "BCTypeNW": self.BCW[i] if j == 0 else None,
                    "BCTypeSE": self.BCN[j] if i == ny - 1 else None,
                    "BCTypeSW": self.BCS[j] if i == 0 else None,
                }
        return out


class RectagularMeshGenerator:
    """Axis-aligned rectangle split into n_blocks_horizontal x n_blocks_vertical blocks
    (mesh/rectangular.py:23-62)."""

    @staticmethod
    def generate(BCE, BCW, BCN, BCS, east: float, west: float, north: float, south: float,
                 n_blocks_horizontal: int, n_blocks_vertical: int) -> QuadMeshGenerator:
        if east <= west:
            raise ValueError(f"East value {east} must be larger than west {west}")
        if north <= south:
            raise ValueError(f"North value {north} must be larger than south {south}")
        tb_y = np.ones(n_blocks_horizontal + 1)
        tb_x = np.linspace(west, east, n_blocks_horizontal + 1)
        ew_x = np.ones(n_blocks_vertical + 1)
        ew_y = np.linspace(south, north, n_blocks_vertical + 1)
        return QuadMeshGenerator(nx_blk=n_blocks_horizontal, ny_blk=n_blocks_vertical, BCE=BCE, BCW=BCW,
                                 BCN=BCN, BCS=BCS, top_x=tb_x, bot_x=tb_x, top_y=north * tb_y,
                                 bot_y=south * tb_y, left_x=west * ew_x, right_x=east * ew_x,
                                 left_y=ew_y, right_y=ew_y)


def mesh_dict_from_block_descriptions(descriptions: dict) -> dict:
    """The mesh dictionary a solver was built from, recovered from the reference's per-block
    BlockDescription objects (blocks/base.py:777-843; Solver.get_mesh_config, solvers/base.py:90-100,
    keeps only these): nBLK from ``info``, the corners from ``geometry.vertices``, neighbours and
    boundary types (sides and corners) from the ``neighbors`` / ``bc`` property dicts."""
    out = {}
    for n, d in descriptions.items():
        v, nb, bc = d.geometry.vertices, d.neighbors, d.bc
        out[n] = {"nBLK": d.info.nBLK, "NE": v.NE, "NW": v.NW, "SE": v.SE, "SW": v.SW,
                  "NeighborE": nb.E, "NeighborW": nb.W, "NeighborN": nb.N, "NeighborS": nb.S,
                  "BCTypeE": bc.E, "BCTypeW": bc.W, "BCTypeN": bc.N, "BCTypeS": bc.S,
                  "BCTypeNE": bc.NE, "BCTypeNW": bc.NW, "BCTypeSE": bc.SE, "BCTypeSW": bc.SW}
    return out


# ------------------------------------------------------------------------------------------------
# geometry of one block
# ------------------------------------------------------------------------------------------------
_libm_sin = np.frompyfunc(math.sin, 1, 1)
_libm_cos = np.frompyfunc(math.cos, 1, 1)


def libm_sincos(theta: np.ndarray):
    return (np.asarray(_libm_sin(theta), dtype=np.float64), np.asarray(_libm_cos(theta), dtype=np.float64))


@dataclass
class FaceSet:
    """Per-cell arrays of one face family (shape (ny, nx)), CellFace + QuadraturePoint."""
    L: np.ndarray
    theta: np.ndarray
    nx: np.ndarray
    ny: np.ndarray
    midx: np.ndarray
    midy: np.ndarray
    qx: np.ndarray
    qy: np.ndarray


class BlockGeometry:
    """Geometry of one (ny, nx) block with corner vertices SW, SE, NW, NE."""

    def __init__(self, vertices: dict, nx: int, ny: int):
        self.nx, self.ny = int(nx), int(ny)
        self.vertices = {k: vertices[k] for k in ("SW", "SE", "NW", "NE")}
        self._build()

    # -------------------------------------------------------------- reference formulas per cell
    def _build(self):
        nx, ny = self.nx, self.ny
        v = self.vertices
        # nodes: per-row linspace between the interpolated west/east edge nodes (quad_mesh.py:78-97)
        e_x = np.linspace(v["SE"][0], v["NE"][0], ny + 1)
        e_y = np.linspace(v["SE"][1], v["NE"][1], ny + 1)
        w_x = np.linspace(v["SW"][0], v["NW"][0], ny + 1)
        w_y = np.linspace(v["SW"][1], v["NW"][1], ny + 1)
        X = np.empty((ny + 1, nx + 1))
        Y = np.empty((ny + 1, nx + 1))
        for i in range(ny + 1):
            X[i] = np.linspace(w_x[i], e_x[i], nx + 1)
            Y[i] = np.linspace(w_y[i], e_y[i], nx + 1)
        self.node_x, self.node_y = X, Y
        ne = (X[1:, 1:], Y[1:, 1:])
        nw = (X[1:, :-1], Y[1:, :-1])
        se = (X[:-1, 1:], Y[:-1, 1:])
        sw = (X[:-1, :-1], Y[:-1, :-1])
        # centroid: 0.25*(((NE + NW) + SE) + SW)  (quad_mesh.py:186-199)
        self.xc = 0.25 * (ne[0] + nw[0] + se[0] + sw[0])
        self.yc = 0.25 * (ne[1] + nw[1] + se[1] + sw[1])

        def face(high, low, vertical: bool, direction: int, q1, q2) -> FaceSet:
            hx, hy = high
            lx, ly = low
            L = np.sqrt((hx - lx) ** 2 + (hy - ly) ** 2)
            with np.errstate(divide="ignore", invalid="ignore"):
                th = np.arctan((lx - hx) / (hy - ly)) if vertical else np.pi / 2 - np.arctan((hy - ly) / (lx - hx))
            # Gauss point p = 0: 0.5*((p2 - p1)*0 + (p2 + p1)) with (p1, p2) = (q1, q2)
            qx = 0.5 * ((q2[0] - q1[0]) * 0 + (q2[0] + q1[0]))
            qy = 0.5 * ((q2[1] - q1[1]) * 0 + (q2[1] + q1[1]))
            return FaceSet(L, th, direction * np.cos(th), direction * np.sin(th),
                           0.5 * (hx + lx), 0.5 * (hy + ly), qx, qy)

        self.face = {
            "E": face(ne, se, True, 1, ne, se),
            "W": face(nw, sw, True, -1, nw, sw),
            "N": face(nw, ne, False, 1, ne, nw),
            "S": face(sw, se, False, -1, se, sw),
        }

        def flen(x1, x2, y1, y2):  # QuadMesh._face_length (mesh/quad_mesh.py:258-268)
            dx = x1 - x2
            dy = y1 - y2
            return np.sqrt(dx * dx + dy * dy)

        s1 = flen(sw[0], nw[0], nw[1], sw[1])
        s3 = flen(se[0], ne[0], ne[1], se[1])
        s2 = flen(nw[0], ne[0], ne[1], nw[1])
        s4 = flen(sw[0], se[0], se[1], sw[1])
        d2 = (X[1:, :-1] - X[:-1, 1:]) ** 2 + (Y[1:, :-1] - Y[:-1, 1:]) ** 2
        with np.errstate(invalid="ignore"):
            a1 = np.arccos((s1 ** 2 + s4 ** 2 - d2) / (2 * s1 * s4))
            a2 = np.arccos((s2 ** 2 + s3 ** 2 - d2) / (2 * s2 * s3))
        s = 0.5 * (s1 + s2 + s3 + s4)
        p1 = (s - s1) * (s - s2) * (s - s3) * (s - s4)
        p2 = s1 * s2 * s3 * s4
        self.A = np.sqrt(p1 - 0.5 * p2 * (1 + np.cos(a1 + a2)))
        self.dx = self.face["E"].midx - self.face["W"].midx
        self.dy = self.face["N"].midy - self.face["S"].midy
        self.is_cartesian = bool((v["NE"][1] == v["NW"][1]) and (v["SE"][1] == v["SW"][1])
                                 and (v["SE"][0] == v["NE"][0]) and (v["SW"][0] == v["NW"][0]))
        if not np.all(np.isfinite(self.A)) or np.any(self.A <= 0):
            raise ValueError("block geometry has degenerate cells")

    # ------------------------------------------------------------------------ derived per cell
    def xl(self, d):
        return self.face[d].L * self.face[d].nx

    def yl(self, d):
        return self.face[d].L * self.face[d].ny

    def boundary_theta(self, d: str) -> np.ndarray:
        """QuadMesh.boundary_angle (mesh/quad_mesh.py:362-369) as a 1-D array along the side."""
        th = self.face[d].theta
        return {"E": th[:, -1], "W": th[:, 0], "N": th[-1, :], "S": th[0, :]}[d]

    # ------------------------------------------------------------------------ packed layouts
    def pack_cartesian(self) -> Optional[np.ndarray]:
        """Separable layout (include/pyhype_hip.h PHX_GEOM_CARTESIAN) or None if the block's
        per-cell arrays are not exactly separable."""
        if not self.is_cartesian:
            return None
        f = self.face

        def colv(a):  # (ny, nx) -> (nx,) if every row is identical
            return a[0].copy() if np.array_equal(a, np.broadcast_to(a[0], a.shape)) else None

        def rowv(a):
            return a[:, 0].copy() if np.array_equal(a, np.broadcast_to(a[:, :1], a.shape)) else None

        cols = [self.xl("N"), self.xl("S"), self.yl("N"), self.yl("S"), f["N"].L, f["S"].L,
                f["E"].qx - self.xc, f["W"].qx - self.xc, f["N"].qx - self.xc, f["S"].qx - self.xc, self.dx]
        rows = [self.xl("E"), self.xl("W"), self.yl("E"), self.yl("W"), f["E"].L, f["W"].L,
                f["E"].qy - self.yc, f["W"].qy - self.yc, f["N"].qy - self.yc, f["S"].qy - self.yc, self.dy]
        c1 = [colv(a) for a in cols]
        r1 = [rowv(a) for a in rows]
        if any(c is None for c in c1) or any(r is None for r in r1):
            return None
        wall = []
        for d in ("E", "W", "N", "S"):
            th = self.boundary_theta(d)
            if not np.all(th == th[0]):
                return None
            sn, cs = libm_sincos(th[:1])
            wall += [cs[0], sn[0]]
        return np.concatenate([np.array(wall), *c1, *r1, self.A.ravel()]).astype(np.float64)

    def pack_general(self) -> np.ndarray:
        """Per-face layout (include/pyhype_hip.h PHX_GEOM_GENERAL), verified against the per-cell
        face arrays (E face of cell j == vertical face j+1, W face == vertical face j with the
        normal negated, N/S likewise)."""
        nx, ny = self.nx, self.ny
        f = self.face
        V = np.empty((7, ny, nx + 1))
        H = np.empty((7, ny + 1, nx))
        # vertical faces: j = 0..nx-1 from the west face of cell j, j = nx from the east face of cell nx-1
        for q, (wq, eq) in enumerate((
                (f["W"].L, f["E"].L), (-self.xl("W"), self.xl("E")), (-self.yl("W"), self.yl("E")),
                (f["W"].qx, f["E"].qx), (f["W"].qy, f["E"].qy))):
            V[q, :, :nx] = wq
            V[q, :, nx] = eq[:, -1]
        thV = np.empty((ny, nx + 1))
        thV[:, :nx] = f["W"].theta
        thV[:, nx] = f["E"].theta[:, -1]
        V[6], V[5] = libm_sincos(thV)
        for q, (sq, nq) in enumerate((
                (f["S"].L, f["N"].L), (-self.xl("S"), self.xl("N")), (-self.yl("S"), self.yl("N")),
                (f["S"].qx, f["N"].qx), (f["S"].qy, f["N"].qy))):
            H[q, :ny, :] = sq
            H[q, ny, :] = nq[-1, :]
        thH = np.empty((ny + 1, nx))
        thH[:ny] = f["S"].theta
        thH[ny] = f["N"].theta[-1]
        H[6], H[5] = libm_sincos(thH)
        # exact consistency of the shared-face representation
        checks = [
            (V[0, :, 1:], f["E"].L), (V[1, :, 1:], self.xl("E")), (V[2, :, 1:], self.yl("E")),
            (V[3, :, 1:], f["E"].qx), (V[4, :, 1:], f["E"].qy), (thV[:, 1:], f["E"].theta),
            (V[0, :, :-1], f["W"].L), (-V[1, :, :-1], self.xl("W")), (-V[2, :, :-1], self.yl("W")),
            (H[0, 1:], f["N"].L), (H[1, 1:], self.xl("N")), (H[2, 1:], self.yl("N")),
            (H[3, 1:], f["N"].qx), (H[4, 1:], f["N"].qy), (thH[1:], f["N"].theta),
            (H[0, :-1], f["S"].L), (-H[1, :-1], self.xl("S")), (-H[2, :-1], self.yl("S")),
        ]
        for a, b in checks:
            if not np.array_equal(a, b):
                raise ValueError("block faces are not consistently shared between neighbouring cells")
        C = np.stack([self.A, self.xc, self.yc, self.dx, self.dy])
        return np.concatenate([V.ravel(), H.ravel(), C.ravel()]).astype(np.float64)

    def pack_quadrature(self, nq: int) -> np.ndarray:
        """Gauss-point offsets from the centroid (include/pyhype_hip.h, after the packed layout):
        for face F in E, W, N, S and point i of the nq-point rule, xq - xc then yq - yc per cell,
        with QuadraturePointData._transform 0.5*((p2 - p1)*p + (p2 + p1)) of the face's vertex pair
        (p1, p2) = E (NE, SE), W (NW, SW), N (NE, NW), S (SE, SW) (mesh/quadratures.py:56-95)."""
        X, Y = self.node_x, self.node_y
        ne, nw, se, sw = (X[1:, 1:], Y[1:, 1:]), (X[1:, :-1], Y[1:, :-1]), (X[:-1, 1:], Y[:-1, 1:]), (X[:-1, :-1], Y[:-1, :-1])
        ends = {"E": (ne, se), "W": (nw, sw), "N": (ne, nw), "S": (se, sw)}
        out = []
        for d in SIDES:
            p1, p2 = ends[d]
            for p, _ in QUADS[nq]:
                qx = 0.5 * ((p2[0] - p1[0]) * p + (p2[0] + p1[0]))
                qy = 0.5 * ((p2[1] - p1[1]) * p + (p2[1] + p1[1]))
                out += [(qx - self.xc).ravel(), (qy - self.yc).ravel()]
        return np.concatenate(out).astype(np.float64)

    def wall_cos_sin(self, d: str):
        sn, cs = libm_sincos(self.boundary_theta(d))
        return cs, sn
