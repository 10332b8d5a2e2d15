// phx_device.hpp -- block descriptors, geometry accessors, ghost values, boundary-face states and
// rotated face fluxes shared by the stage kernels and the adjoint kernels.
//
// Every function is __host__ __device__: the device kernels call them per lane, and the host build
// of the same header runs them in the CPU test harness of the adjoint.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/pyhype_hip.h"
#include "phx_numerics.hpp"

namespace phx {
namespace dev {

using phx::Phys;
using phx::Q4;
using phx::Q4T;

// Device copy of one phx_block_desc.  `off` is the block's offset (in doubles) into the state;
// `wall` holds the (cos, sin) pairs of the E, W, N, S boundary normals of a Cartesian block.
struct DevBlock {
    int nx, ny, case_id, kind;
    int side[4];
    int nbr[4];
    long long off;
    long long yoff;  // offset (doubles) of the block's (ny+1) x nx x 4 y-face fluxes in the split stage
    long long foff;  // offset (faces) of the block's ny*(nx+1) x-faces, then (ny+1)*nx y-faces (element-wise stage)
    const double* geom;
    const double* qoff;  // Gauss-point offsets (8*nq arrays of ny x nx) when nq > 1, else null
    double wall[8];
};

// ---------------------------------------------------------------------------------------------
// memory helpers
// ---------------------------------------------------------------------------------------------
__host__ __device__ inline Q4 ld4(const double* p) {
    const double2* v = reinterpret_cast<const double2*>(p);
    const double2 a = v[0], b = v[1];
    Q4 q;
    q.v[0] = a.x; q.v[1] = a.y; q.v[2] = b.x; q.v[3] = b.y;
    return q;
}
__host__ __device__ inline void st4(double* p, const Q4& q) {
    double2* v = reinterpret_cast<double2*>(p);
    v[0] = make_double2(q.v[0], q.v[1]);
    v[1] = make_double2(q.v[2], q.v[3]);
}

// Element i of an array in device global memory.  The geometry pointers come from the block
// descriptors in memory and carry no address space, so device code states it: the compiler then
// emits global loads instead of flat ones (which also wait on the LDS/scalar counter).
__host__ __device__ inline double gld(const double* p, long i) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(1))) double*)p)[i];
#else
    return p[i];
#endif
}

// Element i of a descriptor table that no kernel writes while it runs (block descriptors, work
// lists, per-case dt).  Device code reads it through the constant address space, so wave-uniform
// reads become scalar loads instead of vector loads that wait behind the state prefetch.
template <class T>
__host__ __device__ inline T ld_const(const T* p, long i) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(4))) T*)p)[i];
#else
    return p[i];
#endif
}

// ---------------------------------------------------------------------------------------------
// geometry accessors (layouts documented in include/pyhype_hip.h)
// ---------------------------------------------------------------------------------------------
// Geometry element i through the constant address space (CG = true; nothing writes the geometry
// while a kernel runs): wave-uniform reads become scalar loads, and the compiler may move the loads
// across the kernel's stores (e.g. out of a row loop).  CG = false: a global load.
template <bool CG>
__host__ __device__ inline double gldg(const double* p, long i) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (CG) return ((const __attribute__((address_space(4))) double*)p)[i];
#endif
    return gld(p, i);
}

template <bool CG = false>
struct CartGT {
    const double* g;
    int nx, ny;
    __host__ __device__ double col(int q, int j) const { return gldg<CG>(g, 8 + q * nx + j); }
    __host__ __device__ double row(int q, int k) const { return gldg<CG>(g, 8 + 11 * nx + q * ny + k); }
    __host__ __device__ double A(int k, int j) const {
        return gldg<CG>(g, 8 + 11 * nx + 11 * ny + (long)k * nx + j);
    }
};
using CartG = CartGT<false>;
enum { C_xlN, C_xlS, C_ylN, C_ylS, C_LN, C_LS, C_dxE, C_dxW, C_dxN, C_dxS, C_dx };
enum { R_xlE, R_xlW, R_ylE, R_ylW, R_LE, R_LW, R_dyE, R_dyW, R_dyN, R_dyS, R_dy };

struct GenG {
    const double* g;
    int nx, ny;
    __host__ __device__ long nV() const { return (long)ny * (nx + 1); }
    __host__ __device__ long nH() const { return (long)(ny + 1) * nx; }
    __host__ __device__ double V(int q, int k, int jf) const { return gld(g, q * nV() + (long)k * (nx + 1) + jf); }
    __host__ __device__ double H(int q, int kf, int j) const { return gld(g, 7 * nV() + q * nH() + (long)kf * nx + j); }
    __host__ __device__ double C(int q, int k, int j) const {
        return gld(g, 7 * nV() + 7 * nH() + q * (long)ny * nx + (long)k * nx + j);
    }
};
enum { F_L, F_xl, F_yl, F_xq, F_yq, F_c, F_s };
enum { CL_A, CL_xc, CL_yc, CL_dx, CL_dy };

// Column arrays of a Cartesian block at column j (they do not depend on the row): loaded once per
// work item by kernels that walk the rows of a strip.
struct CartCols {
    double xlN, xlS, ylN, ylS, dxE, dxW, dxN, dxS, LN, LS, dx;
};
template <bool CG>
__host__ __device__ inline CartCols cart_cols(const DevBlock& B, int j) {
    const CartGT<CG> G{B.geom, B.nx, B.ny};
    CartCols c;
    c.xlN = G.col(C_xlN, j); c.xlS = G.col(C_xlS, j);
    c.ylN = G.col(C_ylN, j); c.ylS = G.col(C_ylS, j);
    c.dxE = G.col(C_dxE, j); c.dxW = G.col(C_dxW, j);
    c.dxN = G.col(C_dxN, j); c.dxS = G.col(C_dxS, j);
    c.LN = G.col(C_LN, j); c.LS = G.col(C_LS, j);
    c.dx = G.col(C_dx, j);
    return c;
}

// cell_geo of a Cartesian cell from its column arrays, the row arrays of row k and its area Ac
template <bool CG>
__host__ __device__ inline phx::CellGeo cart_cell_geo(const DevBlock& B, const CartCols& c, int k, double Ac) {
    const CartGT<CG> G{B.geom, B.nx, B.ny};
    phx::CellGeo g;
    g.xlE = G.row(R_xlE, k); g.xlW = G.row(R_xlW, k);
    g.ylE = G.row(R_ylE, k); g.ylW = G.row(R_ylW, k);
    g.xlN = c.xlN; g.xlS = c.xlS;
    g.ylN = c.ylN; g.ylS = c.ylS;
    g.ainv = 1.0 / Ac;
    g.dxE = c.dxE; g.dxW = c.dxW;
    g.dxN = c.dxN; g.dxS = c.dxS;
    g.dyE = G.row(R_dyE, k); g.dyW = G.row(R_dyW, k);
    g.dyN = G.row(R_dyN, k); g.dyS = G.row(R_dyS, k);
    return g;
}

template <bool CART>
__host__ __device__ inline void wall_cs(const DevBlock& B, int side, int k, int j, double& c, double& s) {
    if (CART) {
        c = B.wall[2 * side];
        s = B.wall[2 * side + 1];
    } else {
        const GenG G{B.geom, B.nx, B.ny};
        switch (side) {
            case PHX_E: c = G.V(F_c, k, B.nx); s = G.V(F_s, k, B.nx); break;
            case PHX_W: c = G.V(F_c, k, 0); s = G.V(F_s, k, 0); break;
            case PHX_N: c = G.H(F_c, B.ny, j); s = G.H(F_s, B.ny, j); break;
            default: c = G.H(F_c, 0, j); s = G.H(F_s, 0, j); break;
        }
    }
}

// Ghost value beyond side SIDE of block B at (k, j): neighbour cell, or own boundary cell
// (reflected for walls).  SIDE is a compile-time constant so that the descriptor fields read here
// (side kind, neighbour block) are wave-uniform scalar loads.
template <bool CART, int SIDE>
__host__ __device__ inline Q4 ghost_value(const DevBlock* __restrict__ blocks, const DevBlock& B,
                                          const double* __restrict__ U, int k, int j) {
    const int st = B.side[SIDE];
    if (st == PHX_SIDE_NEIGHBOR) {
        const DevBlock N = ld_const(blocks, B.nbr[SIDE]);
        const int nk = SIDE == PHX_N ? 0 : (SIDE == PHX_S ? N.ny - 1 : k);
        const int nj = SIDE == PHX_E ? 0 : (SIDE == PHX_W ? N.nx - 1 : j);
        return ld4(U + N.off + 4 * ((long)nk * N.nx + nj));
    }
    const int kk = SIDE == PHX_N ? B.ny - 1 : (SIDE == PHX_S ? 0 : k);
    const int jj = SIDE == PHX_E ? B.nx - 1 : (SIDE == PHX_W ? 0 : j);
    Q4 q = ld4(U + B.off + 4 * ((long)kk * B.nx + jj));
    if (st == PHX_SIDE_WALL) {
        double c, s;
        wall_cs<CART>(B, SIDE, k, j, c, s);
        q = phx::reflect(q, c, s);
    }
    return q;
}

// Value of cell (k, j) of block B, where (k, j) may lie one cell outside the block on one side
// (never a corner): Blocks.apply_boundary_condition / GhostBlock.send_boundary_data
// (blocks/base.py:448-471, blocks/ghost.py:187-216) followed by the ghost's BC (ghost.py:243-278).
template <bool CART>
__host__ __device__ inline Q4 fetch(const DevBlock* __restrict__ blocks, const DevBlock& B, const double* __restrict__ U,
                    int k, int j) {
    const int nx = B.nx, ny = B.ny;
    if (k >= 0 && k < ny && j >= 0 && j < nx) return ld4(U + B.off + 4 * ((long)k * nx + j));
    if (j < 0) return ghost_value<CART, PHX_W>(blocks, B, U, k, j);
    if (j >= nx) return ghost_value<CART, PHX_E>(blocks, B, U, k, j);
    if (k < 0) return ghost_value<CART, PHX_S>(blocks, B, U, k, j);
    return ghost_value<CART, PHX_N>(blocks, B, U, k, j);
}

// fetch() in the reconstruction variables: with GEN and PrimitiveState reconstruction the value
// (for walls the reflected conservative ghost, as the reference's ghost blocks hold it) converted by
// ConservativeConverter.to_primitive (recon_block.from_block, blocks/quad_block.py:120-134).
template <bool CART, bool GEN>
__host__ __device__ inline Q4 fetch_r(const DevBlock* __restrict__ blocks, const DevBlock& B,
                                      const double* __restrict__ U, int k, int j, const Phys& ph) {
    const Q4 q = fetch<CART>(blocks, B, U, k, j);
    if (GEN && ph.prim) return phx::prim_q4(phx::cons_to_prim(q, ph));
    return q;
}

// Outer state of a boundary face: MUSCL._get_boundary_flux_states (fvm/base.py:305-325) and the
// BC applied to it (fvm/base.py:352-362,421-431).  `own` is the block's limited face value,
// `ghost` the ghost cell value (the ghost block's limited value with phi = 0, grad = 0).
template <bool CART, class T>
__host__ __device__ inline Q4T<T> outer_state(const DevBlock& B, int side, const Q4T<T>& own, const Q4T<T>& ghost,
                                              int k, int j) {
    const int st = B.side[side];
    if (st == PHX_SIDE_NEIGHBOR || st == PHX_SIDE_OPEN) return ghost;
    if (st == PHX_SIDE_OUTLET) return own;
    double c, s;
    wall_cs<CART>(B, side, k, j, c, s);
    return phx::reflect(own, c, s);
}

// Face flux of face states in the reconstruction variables; GEN = false: conservative states.
template <int FLUX, bool GEN, class T>
__host__ __device__ inline Q4T<T> flux_of(const Q4T<T>& L, const Q4T<T>& R, const Phys& ph) {
    return GEN ? phx::face_flux_recon<FLUX>(L, R, ph) : phx::face_flux<FLUX>(L, R, ph);
}

template <bool CART, int FLUX, bool GEN = false, class T>
__host__ __device__ inline Q4T<T> xflux(const DevBlock& B, const Q4T<T>& L, const Q4T<T>& R, int k, int jf,
                                        const Phys& ph) {
    if (CART) return flux_of<FLUX, GEN>(L, R, ph);
    const GenG G{B.geom, B.nx, B.ny};
    const double c = G.V(F_c, k, jf), s = G.V(F_s, k, jf);
    const Q4T<T> F = flux_of<FLUX, GEN>(phx::rotate(L, c, s), phx::rotate(R, c, s), ph);
    return phx::unrotate(F, c, s);
}

template <bool CART, int FLUX, bool GEN = false, class T>
__host__ __device__ inline Q4T<T> yflux(const DevBlock& B, const Q4T<T>& L, const Q4T<T>& R, int kf, int j,
                                        const Phys& ph) {
    if (CART) {
        const Q4T<T> F = flux_of<FLUX, GEN>(phx::rotate90(L), phx::rotate90(R), ph);
        return phx::unrotate90(F);
    }
    const GenG G{B.geom, B.nx, B.ny};
    const double c = G.H(F_c, kf, j), s = G.H(F_s, kf, j);
    const Q4T<T> F = flux_of<FLUX, GEN>(phx::rotate(L, c, s), phx::rotate(R, c, s), ph);
    return phx::unrotate(F, c, s);
}

// Gauss-point offsets of cell (k, j): plane ((F*nq + i)*2 + c) of B.qoff for face F, point i,
// component c (layout in include/pyhype_hip.h); points i >= nq are zero.
__host__ __device__ inline phx::QGeo cell_qgeo(const DevBlock& B, int k, int j, int nq) {
    phx::QGeo q;
    const long nc = (long)B.nx * B.ny, c = (long)k * B.nx + j;
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
        for (int i = 0; i < phx::QMAX; ++i) {
            q.dx[f][i] = 0.0;
            q.dy[f][i] = 0.0;
            if (i < nq) {
                q.dx[f][i] = gld(B.qoff, (long)((f * nq + i) * 2) * nc + c);
                q.dy[f][i] = gld(B.qoff, (long)((f * nq + i) * 2 + 1) * nc + c);
            }
        }
    return q;
}

template <bool CART>
__host__ __device__ inline phx::CellGeo cell_geo(const DevBlock& B, int k, int j) {
    phx::CellGeo g;
    if (CART) {
        const CartG G{B.geom, B.nx, B.ny};
        g.xlE = G.row(R_xlE, k); g.xlW = G.row(R_xlW, k);
        g.ylE = G.row(R_ylE, k); g.ylW = G.row(R_ylW, k);
        g.xlN = G.col(C_xlN, j); g.xlS = G.col(C_xlS, j);
        g.ylN = G.col(C_ylN, j); g.ylS = G.col(C_ylS, j);
        g.ainv = 1.0 / G.A(k, j);
        g.dxE = G.col(C_dxE, j); g.dxW = G.col(C_dxW, j);
        g.dxN = G.col(C_dxN, j); g.dxS = G.col(C_dxS, j);
        g.dyE = G.row(R_dyE, k); g.dyW = G.row(R_dyW, k);
        g.dyN = G.row(R_dyN, k); g.dyS = G.row(R_dyS, k);
    } else {
        const GenG G{B.geom, B.nx, B.ny};
        g.xlE = G.V(F_xl, k, j + 1); g.xlW = -G.V(F_xl, k, j);
        g.ylE = G.V(F_yl, k, j + 1); g.ylW = -G.V(F_yl, k, j);
        g.xlN = G.H(F_xl, k + 1, j); g.xlS = -G.H(F_xl, k, j);
        g.ylN = G.H(F_yl, k + 1, j); g.ylS = -G.H(F_yl, k, j);
        g.ainv = 1.0 / G.C(CL_A, k, j);
        const double xc = G.C(CL_xc, k, j), yc = G.C(CL_yc, k, j);
        g.dxE = G.V(F_xq, k, j + 1) - xc; g.dyE = G.V(F_yq, k, j + 1) - yc;
        g.dxW = G.V(F_xq, k, j) - xc; g.dyW = G.V(F_yq, k, j) - yc;
        g.dxN = G.H(F_xq, k + 1, j) - xc; g.dyN = G.H(F_yq, k + 1, j) - yc;
        g.dxS = G.H(F_xq, k, j) - xc; g.dyS = G.H(F_yq, k, j) - yc;
    }
    return g;
}

// face lengths of cell (k, j): west, east, south, north, and its area
template <bool CART>
__host__ __device__ inline void cell_L(const DevBlock& B, int k, int j, double& LW, double& LE, double& LS,
                                       double& LN, double& A) {
    if (CART) {
        const CartG G{B.geom, B.nx, B.ny};
        LW = G.row(R_LW, k); LE = G.row(R_LE, k);
        LS = G.col(C_LS, j); LN = G.col(C_LN, j);
        A = G.A(k, j);
    } else {
        const GenG G{B.geom, B.nx, B.ny};
        LW = G.V(F_L, k, j); LE = G.V(F_L, k, j + 1);
        LS = G.H(F_L, k, j); LN = G.H(F_L, k + 1, j);
        A = G.C(CL_A, k, j);
    }
}

template <bool CART>
__host__ __device__ inline void cell_dxdy(const DevBlock& B, int k, int j, double& dx, double& dy) {
    if (CART) {
        const CartG G{B.geom, B.nx, B.ny};
        dx = G.col(C_dx, j);
        dy = G.row(R_dy, k);
    } else {
        const GenG G{B.geom, B.nx, B.ny};
        dx = G.C(CL_dx, k, j);
        dy = G.C(CL_dy, k, j);
    }
}

}  // namespace dev
}  // namespace phx
