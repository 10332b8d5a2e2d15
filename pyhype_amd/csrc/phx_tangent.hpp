// phx_tangent.hpp -- tangent-linear (forward-mode) element operations of the residual.
//
// The residual's arithmetic (phx_numerics.hpp, phx_device.hpp) is written once for a scalar type;
// here it is instantiated with D, a value and a directional derivative.  The value part of every D
// operation is the double operation of the forward path, so each comparison (limiter branches,
// Harten, |.|, HLLL's wave-speed cases, the 5-point max/min) takes the branch the forward takes,
// and the derivative part is the exact derivative of that branch: dR = (dR/dU) dU at U.  Ties pick
// what the reverse pass picks (first maximum/minimum of the stencil, first minimum of the limiter
// values, sign 0 at |0|), so the tangent and the adjoint linearise the same branch.
//   recon_cell_d   U, dU      -> face states E, W, N, S of a cell and their derivatives (16 + 16)
//   face_flux_d    states     -> derivative of the face flux (4 doubles per face)
// The residual's derivative is then cell_residual (phx_adjoint_ops.hpp) of the flux derivatives,
// the residual being linear in the fluxes.  Midpoint faces (one Gauss point) only.
#pragma once
#include "phx_adjoint_ops.hpp"

namespace phx {
namespace tng {

// value and derivative
struct D {
    double v, d;
    D() = default;
    __host__ __device__ explicit D(double x) : v(x), d(0.0) {}
    __host__ __device__ D(double x, double t) : v(x), d(t) {}
};

__host__ __device__ inline D operator+(D a, D b) { return D(a.v + b.v, a.d + b.d); }
__host__ __device__ inline D operator+(D a, double c) { return D(a.v + c, a.d); }
__host__ __device__ inline D operator+(double c, D a) { return D(c + a.v, a.d); }
__host__ __device__ inline D operator-(D a, D b) { return D(a.v - b.v, a.d - b.d); }
__host__ __device__ inline D operator-(D a, double c) { return D(a.v - c, a.d); }
__host__ __device__ inline D operator-(double c, D a) { return D(c - a.v, -a.d); }
__host__ __device__ inline D operator-(D a) { return D(-a.v, -a.d); }
__host__ __device__ inline D operator*(D a, D b) { return D(a.v * b.v, a.d * b.v + a.v * b.d); }
__host__ __device__ inline D operator*(D a, double c) { return D(a.v * c, a.d * c); }
__host__ __device__ inline D operator*(double c, D a) { return D(c * a.v, c * a.d); }
__host__ __device__ inline D operator/(D a, D b) {
    const double q = a.v / b.v;
    return D(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ inline D operator/(D a, double c) { return D(a.v / c, a.d / c); }
__host__ __device__ inline D operator/(double c, D a) {
    const double q = c / a.v;
    return D(q, -(q * a.d) / a.v);
}
__host__ __device__ inline bool operator<(D a, D b) { return a.v < b.v; }
__host__ __device__ inline bool operator<=(D a, D b) { return a.v <= b.v; }
__host__ __device__ inline bool operator>(D a, D b) { return a.v > b.v; }
__host__ __device__ inline bool operator>=(D a, D b) { return a.v >= b.v; }
__host__ __device__ inline bool operator==(D a, D b) { return a.v == b.v; }
__host__ __device__ inline bool operator!=(D a, D b) { return a.v != b.v; }
__host__ __device__ inline bool operator<(D a, double c) { return a.v < c; }
__host__ __device__ inline bool operator<=(D a, double c) { return a.v <= c; }
__host__ __device__ inline bool operator>(D a, double c) { return a.v > c; }
__host__ __device__ inline bool operator>=(D a, double c) { return a.v >= c; }
__host__ __device__ inline bool operator<(double c, D a) { return c < a.v; }
__host__ __device__ inline bool operator<=(double c, D a) { return c <= a.v; }
__host__ __device__ inline bool operator>(double c, D a) { return c > a.v; }
__host__ __device__ inline bool operator>=(double c, D a) { return c >= a.v; }

__host__ __device__ inline D sqrt(D a) {
    const double r = ::sqrt(a.v);
    return D(r, a.d == 0.0 ? 0.0 : a.d / (2.0 * r));
}
// |a|: derivative sign(a) da with sign(0) = 0 (the reverse pass's convention)
__host__ __device__ inline D fabs(D a) {
    return D(::fabs(a.v), (a.v > 0.0) ? a.d : ((a.v < 0.0) ? -a.d : 0.0));
}
// fmax/fmin of the 5-point stencil: the value of the double maxNum/minNum; on ties the first
// operand, so a chain fmax(fmax(u, uw), ...) selects the first maximum in stencil order
__host__ __device__ inline D fmax(D a, D b) {
    if (a.v != a.v) return b;
    if (b.v != b.v) return a;
    return (b.v > a.v) ? b : a;
}
__host__ __device__ inline D fmin(D a, D b) {
    if (a.v != a.v) return b;
    if (b.v != b.v) return a;
    return (b.v < a.v) ? b : a;
}
// HLLL's Euclidean norm: the value of phx::nrm4 (BLAS dnrm2's arithmetic), derivative x.dx / |x|
__host__ __device__ inline D nrm4(D a, D b, D c, D e) {
    const double n = phx::nrm4(a.v, b.v, c.v, e.v);
    const double t = (n > 0.0) ? (((a.v * a.d + b.v * b.d) + c.v * c.d) + e.v * e.d) / n : 0.0;
    return D(n, t);
}

using QD = Q4T<D>;

__host__ __device__ inline QD make(const Q4& a, const Q4& t) {
    QD q;
    for (int c = 0; c < 4; ++c) q.v[c] = D(a.v[c], t.v[c]);
    return q;
}
__host__ __device__ inline Q4 val(const QD& q) {
    Q4 r;
    for (int c = 0; c < 4; ++c) r.v[c] = q.v[c].v;
    return r;
}
__host__ __device__ inline Q4 der(const QD& q) {
    Q4 r;
    for (int c = 0; c < 4; ++c) r.v[c] = q.v[c].d;
    return r;
}

// dev::fetch_r with its derivative: the ghost value is linear in the state (a copy or a wall
// reflection), so the derivative is the same fetch of dU; Primitive reconstruction converts both.
template <bool CART>
__host__ __device__ inline QD fetch_d(const dev::DevBlock* blocks, const dev::DevBlock& B, const double* U,
                                      const double* dU, int k, int j, const Phys& ph) {
    const QD q = make(dev::fetch<CART>(blocks, B, U, k, j), dev::fetch<CART>(blocks, B, dU, k, j));
    if (ph.prim) return prim_q4(cons_to_prim(q, ph));
    return q;
}

// Face states of cell (k, j) (adj::recon_cell, midpoint faces) and their derivatives: values to
// rb, derivatives to tb (16 doubles each: E, W, N, S).
template <bool CART>
__host__ __device__ inline void recon_cell_d(const dev::DevBlock* blocks, const dev::DevBlock& B, const double* U,
                                             const double* dU, int k, int j, double* rb, double* tb,
                                             const Phys& ph) {
    const QD Uc = fetch_d<CART>(blocks, B, U, dU, k, j, ph);
    const QD Uw = fetch_d<CART>(blocks, B, U, dU, k, j - 1, ph);
    const QD Ue = fetch_d<CART>(blocks, B, U, dU, k, j + 1, ph);
    const QD Us = fetch_d<CART>(blocks, B, U, dU, k - 1, j, ph);
    const QD Un = fetch_d<CART>(blocks, B, U, dU, k + 1, j, ph);
    const CellGeo g = dev::cell_geo<CART>(B, k, j);
    const ReconT<D> r = reconstruct<true>(Uc, Uw, Ue, Us, Un, g, ph.limiter);
    adj::st(rb + adj::RB_E, val(r.E));
    adj::st(rb + adj::RB_W, val(r.W));
    adj::st(rb + adj::RB_N, val(r.N));
    adj::st(rb + adj::RB_S, val(r.S));
    adj::st(tb + adj::RB_E, der(r.E));
    adj::st(tb + adj::RB_W, der(r.W));
    adj::st(tb + adj::RB_N, der(r.N));
    adj::st(tb + adj::RB_S, der(r.S));
}

__host__ __device__ inline QD ld_d(const double* rb, const double* tb, long long cell, int off) {
    return make(adj::ld(rb + adj::RB * cell + off), adj::ld(tb + adj::RB * cell + off));
}

// The global-frame flux of face f and its derivative (adj::face_states + adj::face_flux_global with
// one point): interior faces take the cells' limited states, block-edge faces the outer state of
// their boundary from the own limited state and the ghost value.
template <bool CART, int FLUX>
__host__ __device__ inline QD face_flux_dd(const dev::DevBlock* blocks, const dev::DevBlock& B, const double* U,
                                           const double* dU, const double* rb, const double* tb,
                                           const host::AdjBlock& O, long long f, const Phys& ph) {
    const adj::FaceRef F = adj::face_ref(B, f);
    const int nx = B.nx, ny = B.ny;
    QD L, R;
    if (F.xdir) {
        const int k = F.k, jf = F.j;
        if (jf == 0) {
            const QD own = ld_d(rb, tb, adj::cell_index(O, B, k, 0), adj::RB_W);
            L = dev::outer_state<CART>(B, PHX_W, own, fetch_d<CART>(blocks, B, U, dU, k, -1, ph), k, 0);
        } else {
            L = ld_d(rb, tb, adj::cell_index(O, B, k, jf - 1), adj::RB_E);
        }
        if (jf == nx) {
            const QD own = ld_d(rb, tb, adj::cell_index(O, B, k, nx - 1), adj::RB_E);
            R = dev::outer_state<CART>(B, PHX_E, own, fetch_d<CART>(blocks, B, U, dU, k, nx, ph), k, nx);
        } else {
            R = ld_d(rb, tb, adj::cell_index(O, B, k, jf), adj::RB_W);
        }
        return dev::xflux<CART, FLUX, true>(B, L, R, k, jf, ph);
    }
    const int kf = F.k, j = F.j;
    if (kf == 0) {
        const QD own = ld_d(rb, tb, adj::cell_index(O, B, 0, j), adj::RB_S);
        L = dev::outer_state<CART>(B, PHX_S, own, fetch_d<CART>(blocks, B, U, dU, -1, j, ph), 0, j);
    } else {
        L = ld_d(rb, tb, adj::cell_index(O, B, kf - 1, j), adj::RB_N);
    }
    if (kf == ny) {
        const QD own = ld_d(rb, tb, adj::cell_index(O, B, ny - 1, j), adj::RB_N);
        R = dev::outer_state<CART>(B, PHX_N, own, fetch_d<CART>(blocks, B, U, dU, ny, j, ph), ny - 1, j);
    } else {
        R = ld_d(rb, tb, adj::cell_index(O, B, kf, j), adj::RB_S);
    }
    return dev::yflux<CART, FLUX, true>(B, L, R, kf, j, ph);
}

template <bool CART, int FLUX>
__host__ __device__ inline Q4 face_flux_d(const dev::DevBlock* blocks, const dev::DevBlock& B, const double* U,
                                          const double* dU, const double* rb, const double* tb,
                                          const host::AdjBlock& O, long long f, const Phys& ph) {
    return der(face_flux_dd<CART, FLUX>(blocks, B, U, dU, rb, tb, O, f, ph));
}

}  // namespace tng
}  // namespace phx
