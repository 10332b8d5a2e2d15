// phx_adjoint_ops.hpp -- element operations of the residual and of its vector-Jacobian product.
//
// The residual of a block is evaluated element-wise in three passes: (1) the limited face states of
// every cell, (2) the flux of every face from the two states next to it (block-edge faces take
// the outer state of their boundary), (3) the flux balance of every cell.  The same face-state
// selection drives the reverse pass, so that the adjoint is the transpose of exactly this
// evaluation:
//   recon_cell   U          -> face states E, W, N, S of a cell          (stored, 16 doubles per point)
//   face_adj     kappa      -> adjoints of the left/right state of a face (8 doubles per point)
//   cell_adj     face adj.  -> adjoints of the cell and of its 4 stencil neighbours (20 doubles)
//   gather_adj   20-slots   -> lambda_U of a cell (own slot, neighbours' slots, ghost and
//                              outer-state paths of block edges, neighbour blocks)
// With nq Gauss points per face (NQ = QMAX instantiations, nq = Phys::nq) a cell stores the states
// of every point, (face f, point i) at 4*(f*nq + i), and a face the adjoints of its left and right
// states at every point, point i at 8*i; NQ = 1 is the midpoint rule (nq = 1, the same layout).
// Every function is __host__ __device__; the device kernels run one element per lane and the
// host build runs the same code in the CPU test harness.
#pragma once
#include "phx_adjoint_numerics.hpp"
#include "phx_device.hpp"
#include "phx_plan.hpp"

namespace phx {
namespace adj {

using dev::DevBlock;
using host::AdjBlock;

constexpr int RB = 16;    // doubles per cell of the face-state array: E, W, N, S
constexpr int FA = 8;     // doubles per face of the face-adjoint array: left, right
constexpr int SLOT = 20;  // doubles per cell of the slot array: cell, W, E, S, N

enum { RB_E = 0, RB_W = 4, RB_N = 8, RB_S = 12 };
enum { SL_C = 0, SL_W = 4, SL_E = 8, SL_S = 12, SL_N = 16 };

// Gauss points per face of an instantiation: the midpoint (NQ = 1) or the plan's nq.
template <int NQ>
__host__ __device__ inline int npts(const Phys& ph) {
    return NQ == 1 ? 1 : (ph.nq > 1 ? ph.nq : 1);
}
// offset of the state at point i of face f (E, W, N, S) in a cell's record
__host__ __device__ inline int rbo(int f, int i, int nq) { return 4 * (f * nq + i); }

__host__ __device__ inline long long cell_index(const AdjBlock& O, const DevBlock& B, int k, int j) {
    return O.cell + (long long)k * B.nx + j;
}
__host__ __device__ inline long long xface_index(const AdjBlock& O, const DevBlock& B, int k, int jf) {
    return O.face + (long long)k * (B.nx + 1) + jf;
}
__host__ __device__ inline long long yface_index(const AdjBlock& O, const DevBlock& B, int kf, int j) {
    return O.face + (long long)B.ny * (B.nx + 1) + (long long)kf * B.nx + j;
}

__host__ __device__ inline Q4 ld(const double* p) {
    Q4 q;
    q.v[0] = p[0]; q.v[1] = p[1]; q.v[2] = p[2]; q.v[3] = p[3];
    return q;
}
__host__ __device__ inline void st(double* p, const Q4& q) {
    p[0] = q.v[0]; p[1] = q.v[1]; p[2] = q.v[2]; p[3] = q.v[3];
}
__host__ __device__ inline void acc(Q4& a, const Q4& b) {
    a.v[0] += b.v[0]; a.v[1] += b.v[1]; a.v[2] += b.v[2]; a.v[3] += b.v[3];
}
__host__ __device__ inline Q4 scale(const Q4& a, double c) {
    Q4 r;
    r.v[0] = a.v[0] * c; r.v[1] = a.v[1] * c; r.v[2] = a.v[2] * c; r.v[3] = a.v[3] * c;
    return r;
}
__host__ __device__ inline bool is_zero(const Q4& a) {
    return a.v[0] == 0.0 && a.v[1] == 0.0 && a.v[2] == 0.0 && a.v[3] == 0.0;
}

// ---------------------------------------------------------------------------------------------
// pass 1: limited face states of cell (k, j) (the reconstruction of the stage kernel); the
// variables are independent, and the device runs one lane per cell and variable
// ---------------------------------------------------------------------------------------------
template <bool CART, int NQ = 1>
__host__ __device__ inline void recon_cell(const DevBlock* blocks, const DevBlock& B, const double* U, int k, int j,
                                           double* rb, const Phys& ph) {
    const Q4 Uc = dev::fetch_r<CART, true>(blocks, B, U, k, j, ph);
    const Q4 Uw = dev::fetch_r<CART, true>(blocks, B, U, k, j - 1, ph);
    const Q4 Ue = dev::fetch_r<CART, true>(blocks, B, U, k, j + 1, ph);
    const Q4 Us = dev::fetch_r<CART, true>(blocks, B, U, k - 1, j, ph);
    const Q4 Un = dev::fetch_r<CART, true>(blocks, B, U, k + 1, j, ph);
    const CellGeo g = dev::cell_geo<CART>(B, k, j);
    if (NQ == 1) {
        const Recon r = reconstruct<true>(Uc, Uw, Ue, Us, Un, g, ph.limiter);
        st(rb + RB_E, r.E);
        st(rb + RB_W, r.W);
        st(rb + RB_N, r.N);
        st(rb + RB_S, r.S);
        return;
    }
    const int nq = npts<NQ>(ph);
    const QGeo qg = dev::cell_qgeo(B, k, j, nq);
    const ReconQ r = reconstruct_q(Uc, Uw, Ue, Us, Un, g, qg, nq, ph.limiter);
    for (int f = 0; f < 4; ++f)
        for (int i = 0; i < QMAX; ++i)
            if (i < nq) st(rb + rbo(f, i, nq), qstate(r, qg, f, i));
}

// Component v of a Q4 (selection by constant indices: no private-array indexing on the device).
__host__ __device__ inline double comp(const Q4& q, int v) {
    return v == 0 ? q.v[0] : (v == 1 ? q.v[1] : (v == 2 ? q.v[2] : q.v[3]));
}

// Component v of dev::fetch_r (own conservative cells read directly).
template <bool CART>
__host__ __device__ inline double fetch_var(const DevBlock* blocks, const DevBlock& B, const double* U, int k, int j,
                                            int v, const Phys& ph) {
    if (!ph.prim && k >= 0 && k < B.ny && j >= 0 && j < B.nx) return U[B.off + 4 * ((long long)k * B.nx + j) + v];
    return comp(dev::fetch_r<CART, true>(blocks, B, U, k, j, ph), v);
}

// recon_cell (midpoint faces) for variable v only: the same reconstruct_var on the same values, so
// the four calls v = 0..3 write exactly what recon_cell writes (one lane per cell and variable).
template <bool CART>
__host__ __device__ inline void recon_cell_var(const DevBlock* blocks, const DevBlock& B, const double* U, int k,
                                               int j, int v, double* rb, const Phys& ph) {
    const double u = fetch_var<CART>(blocks, B, U, k, j, v, ph);
    const double uw = fetch_var<CART>(blocks, B, U, k, j - 1, v, ph);
    const double ue = fetch_var<CART>(blocks, B, U, k, j + 1, v, ph);
    const double us = fetch_var<CART>(blocks, B, U, k - 1, j, v, ph);
    const double un = fetch_var<CART>(blocks, B, U, k + 1, j, v, ph);
    const CellGeo g = dev::cell_geo<CART>(B, k, j);
    double qE, qW, qN, qS;
    reconstruct_var<true>(u, uw, ue, us, un, g, ph.limiter, qE, qW, qN, qS);
    rb[RB_E + v] = qE;
    rb[RB_W + v] = qW;
    rb[RB_N + v] = qN;
    rb[RB_S + v] = qS;
}

// ---------------------------------------------------------------------------------------------
// pass 2: the two states of a face.  Face f of block B: x-faces (k, jf), jf = 0..nx, first, then
// y-faces (kf, j), kf = 0..ny.  Interior faces take the limited states of the cells on either
// side; block-edge faces take the outer state of the boundary (dev::outer_state).
// ---------------------------------------------------------------------------------------------
struct FaceRef {
    bool xdir;
    int k, j;  // x-face: row k, face column j;  y-face: face row k, column j
};

// (32-bit index arithmetic: a block has fewer than 2^31 faces)
__host__ __device__ inline FaceRef face_ref(const DevBlock& B, long long f64) {
    FaceRef r;
    int f = (int)f64;
    const int nxf = B.ny * (B.nx + 1);
    if (f < nxf) {
        r.xdir = true;
        r.k = f / (B.nx + 1);
        r.j = f - r.k * (B.nx + 1);
    } else {
        f -= nxf;
        r.xdir = false;
        r.k = f / B.nx;
        r.j = f - r.k * B.nx;
    }
    return r;
}

// The two states at point i of face F (nq points per face; cell records of RB*nq doubles).
template <bool CART>
__host__ __device__ inline void face_states(const DevBlock* blocks, const DevBlock& B, const double* U,
                                            const double* rb, const AdjBlock& O, const FaceRef& F, Q4& L, Q4& R,
                                            const Phys& ph, int nq = 1, int i = 0) {
    const int nx = B.nx, ny = B.ny;
    const long long rs = (long long)RB * nq;
    if (F.xdir) {
        const int k = F.k, jf = F.j;
        if (jf == 0) {
            const Q4 own = ld(rb + rs * cell_index(O, B, k, 0) + rbo(PHX_W, i, nq));
            L = dev::outer_state<CART>(B, PHX_W, own, dev::fetch_r<CART, true>(blocks, B, U, k, -1, ph), k, 0);
        } else {
            L = ld(rb + rs * cell_index(O, B, k, jf - 1) + rbo(PHX_E, i, nq));
        }
        if (jf == nx) {
            const Q4 own = ld(rb + rs * cell_index(O, B, k, nx - 1) + rbo(PHX_E, i, nq));
            R = dev::outer_state<CART>(B, PHX_E, own, dev::fetch_r<CART, true>(blocks, B, U, k, nx, ph), k, nx);
        } else {
            R = ld(rb + rs * cell_index(O, B, k, jf) + rbo(PHX_W, i, nq));
        }
    } else {
        const int kf = F.k, j = F.j;
        if (kf == 0) {
            const Q4 own = ld(rb + rs * cell_index(O, B, 0, j) + rbo(PHX_S, i, nq));
            L = dev::outer_state<CART>(B, PHX_S, own, dev::fetch_r<CART, true>(blocks, B, U, -1, j, ph), 0, j);
        } else {
            L = ld(rb + rs * cell_index(O, B, kf - 1, j) + rbo(PHX_N, i, nq));
        }
        if (kf == ny) {
            const Q4 own = ld(rb + rs * cell_index(O, B, ny - 1, j) + rbo(PHX_N, i, nq));
            R = dev::outer_state<CART>(B, PHX_N, own, dev::fetch_r<CART, true>(blocks, B, U, ny, j, ph), ny - 1, j);
        } else {
            R = ld(rb + rs * cell_index(O, B, kf, j) + rbo(PHX_S, i, nq));
        }
    }
}

// Face flux in the global frame (the forward of pass 2; used by the host residual of the tests):
// the flux at the midpoint, or with nq points the sum of w_i F_i from 0 (integrate_flux,
// fvm/base.py:168-190).
template <bool CART, int FLUX>
__host__ __device__ inline Q4 face_flux_global(const DevBlock* blocks, const DevBlock& B, const double* U,
                                               const double* rb, const AdjBlock& O, long long f, const Phys& ph) {
    const FaceRef F = face_ref(B, f);
    const int nq = ph.nq > 1 ? ph.nq : 1;
    Q4 S = q4zero();
    for (int i = 0; i < nq; ++i) {
        Q4 L, R;
        face_states<CART>(blocks, B, U, rb, O, F, L, R, ph, nq, i);
        const Q4 Fi = F.xdir ? dev::xflux<CART, FLUX, true>(B, L, R, F.k, F.j, ph)
                             : dev::yflux<CART, FLUX, true>(B, L, R, F.k, F.j, ph);
        if (nq == 1) return Fi;
        const double w = gauss_w(nq, i);
        for (int c = 0; c < 4; ++c) S.v[c] = S.v[c] + w * Fi.v[c];
    }
    return S;
}

// Residual of cell (k, j) from the global-frame face fluxes (fvm/base.py:141-190; same association
// as the stage kernels): L*(2F) with the midpoint, L*S with the Gauss sums S.
template <bool CART>
__host__ __device__ inline Q4 cell_residual(const DevBlock& B, const AdjBlock& O, const double* flux, int k, int j,
                                            int nq = 1) {
    double LW, LE, LS, LN, A;
    dev::cell_L<CART>(B, k, j, LW, LE, LS, LN, A);
    const Q4 FW = ld(flux + 4 * xface_index(O, B, k, j));
    const Q4 FE = ld(flux + 4 * xface_index(O, B, k, j + 1));
    const Q4 FS = ld(flux + 4 * yface_index(O, B, k, j));
    const Q4 FN = ld(flux + 4 * yface_index(O, B, k + 1, j));
    Q4 R;
    for (int c = 0; c < 4; ++c) {
        if (nq > 1) {
            const double part = (LW * FW.v[c] - LE * FE.v[c]) + LS * FS.v[c];
            R.v[c] = (0.5 * (part - LN * FN.v[c])) / A;
        } else {
            const double part = (LW * (2.0 * FW.v[c]) - LE * (2.0 * FE.v[c])) + LS * (2.0 * FS.v[c]);
            R.v[c] = (0.5 * (part - LN * (2.0 * FN.v[c]))) / A;
        }
    }
    return R;
}

// Residual seed read from an array in the state layout (Seed of face_adj).
struct PtrSeed {
    const double* p;
    __host__ __device__ Q4 operator()(const DevBlock&, long long cell) const { return ld(p + 4 * cell); }
};

// Adjoint of face f: seed lambda_F = sum over the (up to two) cells of the face of
// kappa_cell * dR_cell/dF, then the flux VJP in the face frame.  Writes left/right adjoints.
// kappa(B, cell) returns the residual seed of a cell.
template <bool CART, int FLUX, class Seed, int NQ = 1>
__host__ __device__ inline void face_adj(const DevBlock* blocks, const DevBlock& B, const double* U, const double* rb,
                                         const AdjBlock& O, const Seed& kappa, long long f, double* out,
                                         const Phys& ph) {
    const FaceRef F = face_ref(B, f);
    Q4 lam = q4zero();
    double LW, LE, LS, LN, A;
    if (F.xdir) {
        if (F.j > 0) {  // east face of cell (k, jf-1): -LE/A
            dev::cell_L<CART>(B, F.k, F.j - 1, LW, LE, LS, LN, A);
            acc(lam, scale(kappa(B, cell_index(O, B, F.k, F.j - 1)), -LE / A));
        }
        if (F.j < B.nx) {  // west face of cell (k, jf): +LW/A
            dev::cell_L<CART>(B, F.k, F.j, LW, LE, LS, LN, A);
            acc(lam, scale(kappa(B, cell_index(O, B, F.k, F.j)), LW / A));
        }
    } else {
        if (F.k > 0) {  // north face of cell (kf-1, j): -LN/A
            dev::cell_L<CART>(B, F.k - 1, F.j, LW, LE, LS, LN, A);
            acc(lam, scale(kappa(B, cell_index(O, B, F.k - 1, F.j)), -LN / A));
        }
        if (F.k < B.ny) {  // south face of cell (kf, j): +LS/A
            dev::cell_L<CART>(B, F.k, F.j, LW, LE, LS, LN, A);
            acc(lam, scale(kappa(B, cell_index(O, B, F.k, F.j)), LS / A));
        }
    }
    // with nq points the residual takes 0.5*L*w_i*F_i/A: the seed of point i is 0.5*w_i times the
    // midpoint rule's (whose 0.5*L*(2F)/A gives L/A)
    const int nq = npts<NQ>(ph);
    for (int i = 0; i < (NQ == 1 ? 1 : QMAX); ++i) {
        if (i >= nq) break;
        const Q4 lam_i = (NQ == 1) ? lam : scale(lam, 0.5 * gauss_w(nq, i));
        Q4 bL = q4zero(), bR = q4zero();
        if (!is_zero(lam_i)) {
            Q4 L, R;
            face_states<CART>(blocks, B, U, rb, O, F, L, R, ph, nq, i);
            if (CART && F.xdir) {
                face_flux_vjp_recon<FLUX>(L, R, lam_i, ph, bL, bR);
            } else if (CART) {
                // F = unrotate90(flux(rotate90 L, rotate90 R)); transposes: rotate90^T = unrotate90
                Q4 bl, br;
                face_flux_vjp_recon<FLUX>(rotate90(L), rotate90(R), rotate90(lam_i), ph, bl, br);
                bL = unrotate90(bl);
                bR = unrotate90(br);
            } else {
                const dev::GenG G{B.geom, B.nx, B.ny};
                double c, s;
                if (F.xdir) { c = G.V(dev::F_c, F.k, F.j); s = G.V(dev::F_s, F.k, F.j); }
                else { c = G.H(dev::F_c, F.k, F.j); s = G.H(dev::F_s, F.k, F.j); }
                // F = unrotate(flux(rotate L, rotate R)); rotate^T = unrotate
                Q4 bl, br;
                face_flux_vjp_recon<FLUX>(rotate(L, c, s), rotate(R, c, s), rotate(lam_i, c, s), ph, bl, br);
                bL = unrotate(bl, c, s);
                bR = unrotate(br, c, s);
            }
        }
        st(out + FA * i, bL);
        st(out + FA * i + 4, bR);
    }
}

// Adjoint of the boundary outer state that is a function of the block's own limited face state
// (outlet: identity; wall: reflection, which is symmetric); zero for neighbour and open sides,
// whose outer state is the ghost cell value.
template <bool CART>
__host__ __device__ inline Q4 outer_own_adj(const DevBlock& B, int side, const Q4& b, int k, int j) {
    const int st_ = B.side[side];
    if (st_ == PHX_SIDE_OUTLET) return b;
    if (st_ == PHX_SIDE_WALL) {
        double c, s;
        dev::wall_cs<CART>(B, side, k, j, c, s);
        return reflect(b, c, s);
    }
    return q4zero();
}

// pass 3 reverse, nq Gauss points per face: adjoints of the 4*nq point states of cell (k, j)
// (face records of FA*nq doubles) -> reconstruction VJP over all points, one variable at a time
// (a boundary point's outer-state term mixes the variables, so it is taken per variable again).
template <bool CART>
__host__ __device__ inline void cell_adj_q(const DevBlock* blocks, const DevBlock& B, const double* U,
                                           const AdjBlock& O, const double* fadj, int k, int j, double* slot,
                                           const Phys& ph) {
    const int nx = B.nx, ny = B.ny;
    const int nq = npts<QMAX>(ph);
    const long long fs = (long long)FA * nq;
    const double* fE = fadj + fs * xface_index(O, B, k, j + 1);      // left states of the east face
    const double* fW = fadj + fs * xface_index(O, B, k, j) + 4;      // right states of the west face
    const double* fN = fadj + fs * yface_index(O, B, k + 1, j);      // left states of the north face
    const double* fS = fadj + fs * yface_index(O, B, k, j) + 4;      // right states of the south face
    for (int q = 0; q < SLOT; ++q) slot[q] = 0.0;
    const Q4 Uc = dev::fetch_r<CART, true>(blocks, B, U, k, j, ph);
    const Q4 Uw = dev::fetch_r<CART, true>(blocks, B, U, k, j - 1, ph);
    const Q4 Ue = dev::fetch_r<CART, true>(blocks, B, U, k, j + 1, ph);
    const Q4 Us = dev::fetch_r<CART, true>(blocks, B, U, k - 1, j, ph);
    const Q4 Un = dev::fetch_r<CART, true>(blocks, B, U, k + 1, j, ph);
    const CellGeo g = dev::cell_geo<CART>(B, k, j);
    const QGeo qg = dev::cell_qgeo(B, k, j, nq);
    for (int v = 0; v < 4; ++v) {
        double bq[4][QMAX];
        bool nz = false;
        for (int i = 0; i < QMAX; ++i) {
            for (int f = 0; f < 4; ++f) bq[f][i] = 0.0;
            if (i >= nq) continue;
            bq[PHX_E][i] = fE[FA * i + v];
            bq[PHX_W][i] = fW[FA * i + v];
            bq[PHX_N][i] = fN[FA * i + v];
            bq[PHX_S][i] = fS[FA * i + v];
            if (j == nx - 1)
                bq[PHX_E][i] += outer_own_adj<CART>(B, PHX_E, ld(fadj + fs * xface_index(O, B, k, nx) + FA * i + 4),
                                                    k, nx).v[v];
            if (j == 0)
                bq[PHX_W][i] += outer_own_adj<CART>(B, PHX_W, ld(fadj + fs * xface_index(O, B, k, 0) + FA * i), k, 0).v[v];
            if (k == ny - 1)
                bq[PHX_N][i] += outer_own_adj<CART>(B, PHX_N, ld(fadj + fs * yface_index(O, B, ny, j) + FA * i + 4),
                                                    ny - 1, j).v[v];
            if (k == 0)
                bq[PHX_S][i] += outer_own_adj<CART>(B, PHX_S, ld(fadj + fs * yface_index(O, B, 0, j) + FA * i), 0, j).v[v];
            for (int f = 0; f < 4; ++f) nz = nz || bq[f][i] != 0.0;
        }
        if (!nz) continue;
        double o[5];
        reconstruct_vjp_var_q(Uc.v[v], Uw.v[v], Ue.v[v], Us.v[v], Un.v[v], g, qg, nq, bq, ph.limiter, o);
        slot[SL_C + v] = o[0];
        slot[SL_W + v] = o[1];
        slot[SL_E + v] = o[2];
        slot[SL_S + v] = o[3];
        slot[SL_N + v] = o[4];
    }
}

// pass 3 reverse: adjoints of the four face states of cell (k, j) -> reconstruction VJP (the
// variables are independent; a variable whose four face adjoints are zero contributes zeros).
template <bool CART, int NQ = 1>
__host__ __device__ inline void cell_adj(const DevBlock* blocks, const DevBlock& B, const double* U,
                                         const AdjBlock& O, const double* fadj, int k, int j, double* slot,
                                         const Phys& ph) {
    if (NQ > 1) {
        cell_adj_q<CART>(blocks, B, U, O, fadj, k, j, slot, ph);
        return;
    }
    const int nx = B.nx, ny = B.ny;
    Q4 bE = ld(fadj + FA * xface_index(O, B, k, j + 1));          // left state of the east face
    Q4 bW = ld(fadj + FA * xface_index(O, B, k, j) + 4);          // right state of the west face
    Q4 bN = ld(fadj + FA * yface_index(O, B, k + 1, j));          // left state of the north face
    Q4 bS = ld(fadj + FA * yface_index(O, B, k, j) + 4);          // right state of the south face
    if (j == nx - 1) acc(bE, outer_own_adj<CART>(B, PHX_E, ld(fadj + FA * xface_index(O, B, k, nx) + 4), k, nx));
    if (j == 0) acc(bW, outer_own_adj<CART>(B, PHX_W, ld(fadj + FA * xface_index(O, B, k, 0)), k, 0));
    if (k == ny - 1) acc(bN, outer_own_adj<CART>(B, PHX_N, ld(fadj + FA * yface_index(O, B, ny, j) + 4), ny - 1, j));
    if (k == 0) acc(bS, outer_own_adj<CART>(B, PHX_S, ld(fadj + FA * yface_index(O, B, 0, j)), 0, j));
    for (int q = 0; q < SLOT; ++q) slot[q] = 0.0;
    if (is_zero(bE) && is_zero(bW) && is_zero(bN) && is_zero(bS)) return;
    const Q4 Uc = dev::fetch_r<CART, true>(blocks, B, U, k, j, ph);
    const Q4 Uw = dev::fetch_r<CART, true>(blocks, B, U, k, j - 1, ph);
    const Q4 Ue = dev::fetch_r<CART, true>(blocks, B, U, k, j + 1, ph);
    const Q4 Us = dev::fetch_r<CART, true>(blocks, B, U, k - 1, j, ph);
    const Q4 Un = dev::fetch_r<CART, true>(blocks, B, U, k + 1, j, ph);
    const CellGeo g = dev::cell_geo<CART>(B, k, j);
    const int lim = ph.limiter;
    for (int v = 0; v < 4; ++v) {
        const double bq[4] = {bE.v[v], bW.v[v], bN.v[v], bS.v[v]};
        if (bq[0] == 0.0 && bq[1] == 0.0 && bq[2] == 0.0 && bq[3] == 0.0) continue;
        double o[5];
        reconstruct_vjp_var(Uc.v[v], Uw.v[v], Ue.v[v], Us.v[v], Un.v[v], g, bq, lim, o);
        slot[SL_C + v] = o[0];
        slot[SL_W + v] = o[1];
        slot[SL_E + v] = o[2];
        slot[SL_S + v] = o[3];
        slot[SL_N + v] = o[4];
    }
}

// cell_adj (midpoint faces) for variable v only: component v of the same face adjoints (outer-state
// terms of block-edge faces taken as Q4 and then selected), the same reconstruct_vjp_var; the four
// calls v = 0..3 write exactly what cell_adj writes.
template <bool CART>
__host__ __device__ inline void cell_adj_var(const DevBlock* blocks, const DevBlock& B, const double* U,
                                             const AdjBlock& O, const double* fadj, int k, int j, int v, double* slot,
                                             const Phys& ph) {
    const int nx = B.nx, ny = B.ny;
    double bq[4] = {fadj[FA * xface_index(O, B, k, j + 1) + v], fadj[FA * xface_index(O, B, k, j) + 4 + v],
                    fadj[FA * yface_index(O, B, k + 1, j) + v], fadj[FA * yface_index(O, B, k, j) + 4 + v]};
    if (j == nx - 1) bq[0] += comp(outer_own_adj<CART>(B, PHX_E, ld(fadj + FA * xface_index(O, B, k, nx) + 4), k, nx), v);
    if (j == 0) bq[1] += comp(outer_own_adj<CART>(B, PHX_W, ld(fadj + FA * xface_index(O, B, k, 0)), k, 0), v);
    if (k == ny - 1)
        bq[2] += comp(outer_own_adj<CART>(B, PHX_N, ld(fadj + FA * yface_index(O, B, ny, j) + 4), ny - 1, j), v);
    if (k == 0) bq[3] += comp(outer_own_adj<CART>(B, PHX_S, ld(fadj + FA * yface_index(O, B, 0, j)), 0, j), v);
    double o[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (bq[0] != 0.0 || bq[1] != 0.0 || bq[2] != 0.0 || bq[3] != 0.0) {
        const double u = fetch_var<CART>(blocks, B, U, k, j, v, ph);
        const double uw = fetch_var<CART>(blocks, B, U, k, j - 1, v, ph);
        const double ue = fetch_var<CART>(blocks, B, U, k, j + 1, v, ph);
        const double us = fetch_var<CART>(blocks, B, U, k - 1, j, v, ph);
        const double un = fetch_var<CART>(blocks, B, U, k + 1, j, v, ph);
        const CellGeo g = dev::cell_geo<CART>(B, k, j);
        reconstruct_vjp_var(u, uw, ue, us, un, g, bq, ph.limiter, o);
    }
    slot[SL_C + v] = o[0];
    slot[SL_W + v] = o[1];
    slot[SL_E + v] = o[2];
    slot[SL_S + v] = o[3];
    slot[SL_N + v] = o[4];
}

// Adjoint, with respect to the edge cell's conservative state q, of the ghost value of a wall side,
// R(q) reflected in conservative form and then taken to the reconstruction variables: b is the
// adjoint of that ghost value.  Reflection is symmetric, so its transpose is itself.
template <bool CART>
__host__ __device__ inline Q4 wall_ghost_adj(const DevBlock& B, int side, const Q4& q, const Q4& b, int k, int j,
                                             const Phys& ph) {
    double c, s;
    dev::wall_cs<CART>(B, side, k, j, c, s);
    if (!ph.prim) return reflect(b, c, s);
    const PrimBar pb{b.v[0], b.v[1], b.v[2], b.v[3]};
    return reflect(cons_to_prim_vjp(reflect(q, c, s), pb, ph), c, s);
}

// lambda_U of cell (k, j): every path by which U(k, j) enters the residual.  Paths through the
// cell's own value in the reconstruction variables (its stencil uses, neighbour blocks' ghosts and
// outer states, open and outlet ghosts, which copy it) are summed first and taken back through the
// conversion to primitive variables once; a wall ghost is the reflected state converted, and is
// taken back on its own.
template <bool CART, int NQ = 1>
__host__ __device__ inline Q4 gather_adj(const DevBlock* blocks, const AdjBlock* offs, int b, int k, int j,
                                         const double* slot, const double* fadj, const double* U, const Phys& ph) {
    const DevBlock& B = blocks[b];
    const AdjBlock& O = offs[b];
    const int nx = B.nx, ny = B.ny;
    const int nq = npts<NQ>(ph);
    // adjoint of a boundary face's outer state that is a ghost cell value, summed over the face's
    // points (the ghost value is the same at every point); sub = 0 left state, 4 right state
    auto fsum = [&](long long face, int sub) {
        Q4 r = ld(fadj + (long long)FA * nq * face + sub);
        for (int i = 1; i < (NQ == 1 ? 1 : QMAX); ++i)
            if (i < nq) acc(r, ld(fadj + (long long)FA * nq * face + FA * i + sub));
        return r;
    };
    const Q4 q = ld(U + B.off + 4LL * ((long long)k * nx + j));
    Q4 own = ld(slot + SLOT * cell_index(O, B, k, j) + SL_C);
    Q4 walls = q4zero();
    // stencil neighbour of cells of the same block
    if (j > 0) acc(own, ld(slot + SLOT * cell_index(O, B, k, j - 1) + SL_E));
    if (j < nx - 1) acc(own, ld(slot + SLOT * cell_index(O, B, k, j + 1) + SL_W));
    if (k > 0) acc(own, ld(slot + SLOT * cell_index(O, B, k - 1, j) + SL_N));
    if (k < ny - 1) acc(own, ld(slot + SLOT * cell_index(O, B, k + 1, j) + SL_S));
    // block edges: a neighbour's ghost (its stencil and its outer state), or this block's own ghost
    auto edge = [&](int side, const Q4& ghost_slot, int gk, int gj, const Q4& open_outer) {
        const int st_ = B.side[side];
        if (st_ == PHX_SIDE_WALL) acc(walls, wall_ghost_adj<CART>(B, side, q, ghost_slot, gk, gj, ph));
        else if (st_ == PHX_SIDE_OUTLET) acc(own, ghost_slot);
        else if (st_ == PHX_SIDE_OPEN) { acc(own, ghost_slot); acc(own, open_outer); }
    };
    if (j == 0) {
        if (B.side[PHX_W] == PHX_SIDE_NEIGHBOR) {
            // this cell is the east ghost of the west neighbour: its stencil and its east outer state
            const int n = B.nbr[PHX_W];
            const DevBlock& N = blocks[n];
            acc(own, ld(slot + SLOT * cell_index(offs[n], N, k, N.nx - 1) + SL_E));
            acc(own, fsum(xface_index(offs[n], N, k, N.nx), 4));
        } else {
            edge(PHX_W, ld(slot + SLOT * cell_index(O, B, k, 0) + SL_W), k, -1, fsum(xface_index(O, B, k, 0), 0));
        }
    }
    if (j == nx - 1) {
        if (B.side[PHX_E] == PHX_SIDE_NEIGHBOR) {
            const int n = B.nbr[PHX_E];
            const DevBlock& N = blocks[n];
            acc(own, ld(slot + SLOT * cell_index(offs[n], N, k, 0) + SL_W));
            acc(own, fsum(xface_index(offs[n], N, k, 0), 0));
        } else {
            edge(PHX_E, ld(slot + SLOT * cell_index(O, B, k, nx - 1) + SL_E), k, nx, fsum(xface_index(O, B, k, nx), 4));
        }
    }
    if (k == 0) {
        if (B.side[PHX_S] == PHX_SIDE_NEIGHBOR) {
            const int n = B.nbr[PHX_S];
            const DevBlock& N = blocks[n];
            acc(own, ld(slot + SLOT * cell_index(offs[n], N, N.ny - 1, j) + SL_N));
            acc(own, fsum(yface_index(offs[n], N, N.ny, j), 4));
        } else {
            edge(PHX_S, ld(slot + SLOT * cell_index(O, B, 0, j) + SL_S), -1, j, fsum(yface_index(O, B, 0, j), 0));
        }
    }
    if (k == ny - 1) {
        if (B.side[PHX_N] == PHX_SIDE_NEIGHBOR) {
            const int n = B.nbr[PHX_N];
            const DevBlock& N = blocks[n];
            acc(own, ld(slot + SLOT * cell_index(offs[n], N, 0, j) + SL_S));
            acc(own, fsum(yface_index(offs[n], N, 0, j), 0));
        } else {
            edge(PHX_N, ld(slot + SLOT * cell_index(O, B, ny - 1, j) + SL_N), ny, j, fsum(yface_index(O, B, ny, j), 4));
        }
    }
    Q4 lam = own;
    if (ph.prim) lam = cons_to_prim_vjp(q, PrimBar{own.v[0], own.v[1], own.v[2], own.v[3]}, ph);
    acc(lam, walls);
    return lam;
}

}  // namespace adj
}  // namespace phx
