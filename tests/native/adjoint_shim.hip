// Host build of the element operations of the residual and of its adjoint
// (pyhype_amd/csrc/phx_adjoint_ops.hpp): the same functions the adjoint kernels run per lane, looped
// on the CPU.  Test infrastructure only: the tests compare (a) the element-wise host residual with
// the oracle, (b) the host adjoint with central differences of that residual, (c) the GPU adjoint
// with the host adjoint.
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../pyhype_amd/csrc/phx_adjoint_ops.hpp"
#include "../../pyhype_amd/csrc/phx_tangent.hpp"

using namespace phx;
using phx::dev::DevBlock;
using phx::host::AdjBlock;

namespace {

// p: gamma, gm1, 1/gm1, gamma/gm1, CFL, limiter id, primitive reconstruction (0/1), Gauss points
Phys make_phys(const double* p) {
    return Phys{p[0], p[1], p[2], p[3], p[4], (int)p[5], (int)p[6], p[7] > 1.0 ? (int)p[7] : 1};
}
Q4 q4(const double* p) { Q4 q; for (int k = 0; k < 4; ++k) q.v[k] = p[k]; return q; }
void put(double* p, const Q4& q) { for (int k = 0; k < 4; ++k) p[k] = q.v[k]; }

struct HostPlan {
    std::vector<DevBlock> blocks;
    std::vector<AdjBlock> offs;
    long long cells = 0, faces = 0, len = 0;
};

HostPlan make_plan(const phx_block_desc* d, int nb, int nq) {
    HostPlan P;
    P.blocks.resize(nb);
    P.offs.resize(nb);
    for (int b = 0; b < nb; ++b) {
        DevBlock& h = P.blocks[b];
        memset(&h, 0, sizeof(h));
        h.nx = d[b].nx; h.ny = d[b].ny; h.case_id = d[b].case_id; h.kind = d[b].geom_kind;
        for (int s = 0; s < 4; ++s) { h.side[s] = d[b].side_type[s]; h.nbr[s] = d[b].neighbor[s]; }
        h.off = d[b].state_offset;
        h.geom = d[b].geom;
        h.qoff = nq > 1 ? d[b].geom + (d[b].geom_len - 8LL * nq * h.nx * h.ny) : nullptr;
        if (h.kind == PHX_GEOM_CARTESIAN)
            for (int q = 0; q < 8; ++q) h.wall[q] = d[b].geom[q];
        P.offs[b].cell = d[b].state_offset / 4;
        P.offs[b].face = P.faces;
        P.faces += (long long)h.ny * (h.nx + 1) + (long long)(h.ny + 1) * h.nx;
        P.len = std::max<long long>(P.len, d[b].state_offset + 4LL * h.nx * h.ny);
    }
    P.cells = P.len / 4;
    return P;
}

// by_var: the midpoint rule through recon_cell_var, one call per variable (as the device runs it)
template <int NQ>
void recon_all(const HostPlan& P, const double* U, std::vector<double>& rb, const Phys& ph, bool by_var = false) {
    const int nq = adj::npts<NQ>(ph);
    rb.assign((size_t)P.cells * adj::RB * nq, 0.0);
    for (size_t b = 0; b < P.blocks.size(); ++b) {
        const DevBlock& B = P.blocks[b];
        const bool cart = B.kind == PHX_GEOM_CARTESIAN;
        for (int k = 0; k < B.ny; ++k)
            for (int j = 0; j < B.nx; ++j) {
                double* out = rb.data() + (long long)adj::RB * nq * adj::cell_index(P.offs[b], B, k, j);
                if (NQ == 1 && by_var) {
                    for (int v = 0; v < 4; ++v) {
                        if (cart) adj::recon_cell_var<true>(P.blocks.data(), B, U, k, j, v, out, ph);
                        else adj::recon_cell_var<false>(P.blocks.data(), B, U, k, j, v, out, ph);
                    }
                } else if (cart) {
                    adj::recon_cell<true, NQ>(P.blocks.data(), B, U, k, j, out, ph);
                } else {
                    adj::recon_cell<false, NQ>(P.blocks.data(), B, U, k, j, out, ph);
                }
            }
    }
}

template <int FLUX, int NQ>
void residual_t(const HostPlan& P, const double* U, double* R, const Phys& ph) {
    std::vector<double> rb, flux((size_t)P.faces * 4);
    recon_all<NQ>(P, U, rb, ph);
    for (size_t b = 0; b < P.blocks.size(); ++b) {
        const DevBlock& B = P.blocks[b];
        const long long nf = (long long)B.ny * (B.nx + 1) + (long long)(B.ny + 1) * B.nx;
        for (long long f = 0; f < nf; ++f) {
            const Q4 F = (B.kind == PHX_GEOM_CARTESIAN)
                             ? adj::face_flux_global<true, FLUX>(P.blocks.data(), B, U, rb.data(), P.offs[b], f, ph)
                             : adj::face_flux_global<false, FLUX>(P.blocks.data(), B, U, rb.data(), P.offs[b], f, ph);
            put(flux.data() + 4 * (P.offs[b].face + f), F);
        }
        for (int k = 0; k < B.ny; ++k)
            for (int j = 0; j < B.nx; ++j) {
                const int nq = adj::npts<NQ>(ph);
                const Q4 r = (B.kind == PHX_GEOM_CARTESIAN)
                                 ? adj::cell_residual<true>(B, P.offs[b], flux.data(), k, j, nq)
                                 : adj::cell_residual<false>(B, P.offs[b], flux.data(), k, j, nq);
                put(R + B.off + 4LL * ((long long)k * B.nx + j), r);
            }
    }
}

template <int FLUX, int NQ>
void residual_vjp_t(const HostPlan& P, const double* U, const double* kappa, double* lamU, const Phys& ph,
                    bool by_var = false) {
    const int nq = adj::npts<NQ>(ph);
    std::vector<double> rb, fadj((size_t)P.faces * adj::FA * nq), slot((size_t)P.cells * adj::SLOT, 0.0);
    recon_all<NQ>(P, U, rb, ph, by_var);
    for (size_t b = 0; b < P.blocks.size(); ++b) {
        const DevBlock& B = P.blocks[b];
        const long long nf = (long long)B.ny * (B.nx + 1) + (long long)(B.ny + 1) * B.nx;
        for (long long f = 0; f < nf; ++f) {
            double* out = fadj.data() + (long long)adj::FA * nq * (P.offs[b].face + f);
            if (B.kind == PHX_GEOM_CARTESIAN)
                adj::face_adj<true, FLUX, adj::PtrSeed, NQ>(P.blocks.data(), B, U, rb.data(), P.offs[b],
                                                            adj::PtrSeed{kappa}, f, out, ph);
            else
                adj::face_adj<false, FLUX, adj::PtrSeed, NQ>(P.blocks.data(), B, U, rb.data(), P.offs[b],
                                                             adj::PtrSeed{kappa}, f, out, ph);
        }
    }
    for (size_t b = 0; b < P.blocks.size(); ++b) {
        const DevBlock& B = P.blocks[b];
        for (int k = 0; k < B.ny; ++k)
            for (int j = 0; j < B.nx; ++j) {
                double* out = slot.data() + adj::SLOT * adj::cell_index(P.offs[b], B, k, j);
                if (NQ == 1 && by_var) {
                    for (int v = 0; v < 4; ++v) {
                        if (B.kind == PHX_GEOM_CARTESIAN)
                            adj::cell_adj_var<true>(P.blocks.data(), B, U, P.offs[b], fadj.data(), k, j, v, out, ph);
                        else
                            adj::cell_adj_var<false>(P.blocks.data(), B, U, P.offs[b], fadj.data(), k, j, v, out, ph);
                    }
                } else if (B.kind == PHX_GEOM_CARTESIAN) {
                    adj::cell_adj<true, NQ>(P.blocks.data(), B, U, P.offs[b], fadj.data(), k, j, out, ph);
                } else {
                    adj::cell_adj<false, NQ>(P.blocks.data(), B, U, P.offs[b], fadj.data(), k, j, out, ph);
                }
            }
    }
    for (size_t b = 0; b < P.blocks.size(); ++b) {
        const DevBlock& B = P.blocks[b];
        for (int k = 0; k < B.ny; ++k)
            for (int j = 0; j < B.nx; ++j) {
                const Q4 lam = (B.kind == PHX_GEOM_CARTESIAN)
                                   ? adj::gather_adj<true, NQ>(P.blocks.data(), P.offs.data(), (int)b, k, j, slot.data(),
                                                               fadj.data(), U, ph)
                                   : adj::gather_adj<false, NQ>(P.blocks.data(), P.offs.data(), (int)b, k, j, slot.data(),
                                                                fadj.data(), U, ph);
                put(lamU + B.off + 4LL * ((long long)k * B.nx + j), lam);
            }
    }
}

// residual through the tangent-linear element operations: values (Rv) and derivatives (dR)
template <int FLUX>
void residual_jvp_t(const HostPlan& P, const double* U, const double* dU, double* Rv, double* dR, const Phys& ph) {
    std::vector<double> rb((size_t)P.cells * adj::RB), tb((size_t)P.cells * adj::RB);
    std::vector<double> Fv((size_t)P.faces * 4), dF((size_t)P.faces * 4);
    for (size_t b = 0; b < P.blocks.size(); ++b) {
        const DevBlock& B = P.blocks[b];
        for (int k = 0; k < B.ny; ++k)
            for (int j = 0; j < B.nx; ++j) {
                const long long c = adj::RB * adj::cell_index(P.offs[b], B, k, j);
                if (B.kind == PHX_GEOM_CARTESIAN)
                    tng::recon_cell_d<true>(P.blocks.data(), B, U, dU, k, j, rb.data() + c, tb.data() + c, ph);
                else
                    tng::recon_cell_d<false>(P.blocks.data(), B, U, dU, k, j, rb.data() + c, tb.data() + c, ph);
            }
    }
    for (size_t b = 0; b < P.blocks.size(); ++b) {
        const DevBlock& B = P.blocks[b];
        const long long nf = (long long)B.ny * (B.nx + 1) + (long long)(B.ny + 1) * B.nx;
        for (long long f = 0; f < nf; ++f) {
            const tng::QD F = (B.kind == PHX_GEOM_CARTESIAN)
                                  ? tng::face_flux_dd<true, FLUX>(P.blocks.data(), B, U, dU, rb.data(), tb.data(),
                                                                  P.offs[b], f, ph)
                                  : tng::face_flux_dd<false, FLUX>(P.blocks.data(), B, U, dU, rb.data(), tb.data(),
                                                                   P.offs[b], f, ph);
            put(Fv.data() + 4 * (P.offs[b].face + f), tng::val(F));
            put(dF.data() + 4 * (P.offs[b].face + f), tng::der(F));
        }
    }
    for (size_t b = 0; b < P.blocks.size(); ++b) {
        const DevBlock& B = P.blocks[b];
        for (int k = 0; k < B.ny; ++k)
            for (int j = 0; j < B.nx; ++j) {
                const bool cart = B.kind == PHX_GEOM_CARTESIAN;
                const long long o = B.off + 4LL * ((long long)k * B.nx + j);
                put(Rv + o, cart ? adj::cell_residual<true>(B, P.offs[b], Fv.data(), k, j)
                                 : adj::cell_residual<false>(B, P.offs[b], Fv.data(), k, j));
                put(dR + o, cart ? adj::cell_residual<true>(B, P.offs[b], dF.data(), k, j)
                                 : adj::cell_residual<false>(B, P.offs[b], dF.data(), k, j));
            }
    }
}

}  // namespace

extern "C" {

// residual (values) and its forward-mode derivative in direction dU (midpoint faces)
int shim_residual_jvp(const phx_block_desc* d, int nb, int flux, const double* phys, const double* U,
                      const double* dU, double* Rv, double* dR) {
    const Phys ph = make_phys(phys);
    if (ph.nq > 1) return -1;
    const HostPlan P = make_plan(d, nb, 1);
    if (flux == 0) residual_jvp_t<0>(P, U, dU, Rv, dR, ph);
    else residual_jvp_t<1>(P, U, dU, Rv, dR, ph);
    return 0;
}

// element-wise residual of every block (same layout as phx_residual)
int shim_residual(const phx_block_desc* d, int nb, int flux, const double* phys, const double* U, double* R) {
    const Phys ph = make_phys(phys);
    const HostPlan P = make_plan(d, nb, ph.nq);
    if (ph.nq > 1) {
        if (flux == 0) residual_t<0, QMAX>(P, U, R, ph);
        else residual_t<1, QMAX>(P, U, R, ph);
    } else {
        if (flux == 0) residual_t<0, 1>(P, U, R, ph);
        else residual_t<1, 1>(P, U, R, ph);
    }
    return 0;
}

// limited face states of every cell (same layout as phx_face_states); by_var: through recon_cell_var
int shim_face_states(const phx_block_desc* d, int nb, const double* phys, const double* U, double* out, int by_var) {
    const Phys ph = make_phys(phys);
    const HostPlan P = make_plan(d, nb, ph.nq);
    std::vector<double> rb;
    if (ph.nq > 1) recon_all<QMAX>(P, U, rb, ph);
    else recon_all<1>(P, U, rb, ph, by_var != 0);
    std::copy(rb.begin(), rb.end(), out);
    return 0;
}

// shim_residual_vjp through the per-variable reconstruction and cell adjoint (midpoint faces)
int shim_residual_vjp_var(const phx_block_desc* d, int nb, int flux, const double* phys, const double* U,
                          const double* kappa, double* lamU) {
    const Phys ph = make_phys(phys);
    if (ph.nq > 1) return -1;
    const HostPlan P = make_plan(d, nb, 1);
    if (flux == 0) residual_vjp_t<0, 1>(P, U, kappa, lamU, ph, true);
    else residual_vjp_t<1, 1>(P, U, kappa, lamU, ph, true);
    return 0;
}

// (dR/dU)^T kappa (same layout as phx_residual_vjp)
int shim_residual_vjp(const phx_block_desc* d, int nb, int flux, const double* phys, const double* U,
                      const double* kappa, double* lamU) {
    const Phys ph = make_phys(phys);
    const HostPlan P = make_plan(d, nb, ph.nq);
    if (ph.nq > 1) {
        if (flux == 0) residual_vjp_t<0, QMAX>(P, U, kappa, lamU, ph);
        else residual_vjp_t<1, QMAX>(P, U, kappa, lamU, ph);
    } else {
        if (flux == 0) residual_vjp_t<0, 1>(P, U, kappa, lamU, ph);
        else residual_vjp_t<1, 1>(P, U, kappa, lamU, ph);
    }
    return 0;
}

// conservative face states -> flux (n faces); and its VJP
void shim_face_flux(int flux, int n, const double* qL, const double* qR, double* F, const double* phys) {
    const Phys ph = make_phys(phys);
    for (int i = 0; i < n; ++i)
        put(F + 4 * i, flux == 0 ? face_flux<0>(q4(qL + 4 * i), q4(qR + 4 * i), ph)
                                 : face_flux<1>(q4(qL + 4 * i), q4(qR + 4 * i), ph));
}
void shim_face_flux_vjp(int flux, int n, const double* qL, const double* qR, const double* lam, double* bL, double* bR,
                        const double* phys) {
    const Phys ph = make_phys(phys);
    for (int i = 0; i < n; ++i) {
        Q4 a, b;
        if (flux == 0) face_flux_vjp<0>(q4(qL + 4 * i), q4(qR + 4 * i), q4(lam + 4 * i), ph, a, b);
        else face_flux_vjp<1>(q4(qL + 4 * i), q4(qR + 4 * i), q4(lam + 4 * i), ph, a, b);
        put(bL + 4 * i, a);
        put(bR + 4 * i, b);
    }
}

// one cell: U5 = (cell, W, E, S, N) x 4, g = 17 doubles in CellGeo order, bQ = (E, W, N, S) x 4,
// out = (cell, W, E, S, N) x 4
void shim_reconstruct_vjp(const double* U5, const double* g, const double* bQ, double* out, int lim) {
    CellGeo G{g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], g[8], g[9], g[10], g[11], g[12], g[13], g[14], g[15], g[16]};
    Recon b;
    b.E = q4(bQ); b.W = q4(bQ + 4); b.N = q4(bQ + 8); b.S = q4(bQ + 12);
    const ReconBar r = reconstruct_vjp(q4(U5), q4(U5 + 4), q4(U5 + 8), q4(U5 + 12), q4(U5 + 16), G, b, lim);
    put(out, r.c); put(out + 4, r.w); put(out + 8, r.e); put(out + 12, r.s); put(out + 16, r.n);
}
void shim_reconstruct1(const double* U5, const double* g, double* out, int lim) {
    CellGeo G{g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], g[8], g[9], g[10], g[11], g[12], g[13], g[14], g[15], g[16]};
    const Recon r = reconstruct<true>(q4(U5), q4(U5 + 4), q4(U5 + 8), q4(U5 + 12), q4(U5 + 16), G, lim);
    put(out, r.E); put(out + 4, r.W); put(out + 8, r.N); put(out + 12, r.S);
}
}
