// phx_adjoint.hip -- discrete adjoint of the MUSCL residual and of one explicit RK step
// (vector-Jacobian products), gfx950 kernels and their C ABI (include/pyhype_hip.h).
//
// The reverse pass is deterministic: every adjoint value is produced by exactly one lane from a
// fixed set of terms (no atomics).  Four element-wise passes per residual VJP, each one lane per
// cell or face, over launch tiles of at most TILE elements of one block:
//   k_adj_recon   limited face states of every cell at the linearisation point
//   k_adj_face    seed of every face from kappa, flux VJP -> left/right state adjoints
//   k_adj_cell    reconstruction VJP of every cell -> cell + 4 stencil-neighbour adjoints
//   (midpoint faces: k_adj_recon_v and k_adj_cell_v, one lane per cell and variable)
//   k_adj_gather  lambda_U of every cell from its own and its neighbours' slots, ghosts, neighbours
// The forward trajectory is not stored by these kernels: the caller passes the stage states,
// recorded by phx_run_record (frozen time steps) from a checkpoint.
#include <hip/hip_runtime.h>

#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/pyhype_hip.h"
#include "phx_adjoint_ops.hpp"
#include "phx_plan.hpp"
#include "phx_tangent.hpp"

using phx::Phys;
using phx::Q4;
using phx::dev::DevBlock;
using phx::dev::ld_const;
using phx::host::AdjBlock;
using phx::host::AdjTile;
using phx::host::fail;

namespace {

constexpr int TILE = 256;
constexpr int MAX_TERMS = PHX_MAX_STAGES + 2;

// sum_m c_m * (dt of the cell's case if scaled_m) * in_m at one cell: the residual seeds and the
// output terms of a Runge-Kutta step VJP, evaluated where they are used
struct Terms {
    const double* in[MAX_TERMS];
    double c[MAX_TERMS];
    int scaled[MAX_TERMS];
    int n;
    const double* dt;
};

__device__ __forceinline__ Q4 eval_terms(const Terms& T, const DevBlock& B, long long cell) {
    Q4 r = phx::q4zero();
    const double dt = T.dt ? T.dt[B.case_id] : 1.0;
#pragma unroll
    for (int m = 0; m < MAX_TERMS; ++m) {
        if (m >= T.n) break;
        const double c = T.scaled[m] ? T.c[m] * dt : T.c[m];
        const double* p = T.in[m] + 4 * cell;
#pragma unroll
        for (int q = 0; q < 4; ++q) r.v[q] += c * p[q];
    }
    return r;
}

struct TermSeed {
    Terms t;
    __device__ Q4 operator()(const DevBlock& B, long long cell) const { return eval_terms(t, B, cell); }
};

// NQ: 1 (midpoint faces) or phx::QMAX (nq = ph.nq Gauss points per face); the point records are
// RB*nq doubles per cell and FA*nq per face
template <int FLUX, int NQ>
__global__ void __launch_bounds__(TILE) k_adj_face(const DevBlock* blocks, const AdjBlock* offs, const AdjTile* tiles,
                                                   const double* U, const double* rb, Terms kappa, double* fadj,
                                                   Phys ph) {
    const AdjTile t = ld_const(tiles, blockIdx.x);
    if ((int)threadIdx.x >= t.count) return;
    const DevBlock B = ld_const(blocks, t.block);  // descriptor in scalar registers
    const AdjBlock O = ld_const(offs, t.block);
    const long long f = (long long)t.start + threadIdx.x;
    double* out = fadj + (long long)phx::adj::FA * phx::adj::npts<NQ>(ph) * (O.face + f);
    const TermSeed seed{kappa};
    if (B.kind == PHX_GEOM_CARTESIAN) phx::adj::face_adj<true, FLUX, TermSeed, NQ>(blocks, B, U, rb, O, seed, f, out, ph);
    else phx::adj::face_adj<false, FLUX, TermSeed, NQ>(blocks, B, U, rb, O, seed, f, out, ph);
}

template <int NQ>
__global__ void __launch_bounds__(TILE) k_adj_recon(const DevBlock* blocks, const AdjBlock* offs, const AdjTile* tiles,
                                                    const double* U, double* rb, Phys ph) {
    const AdjTile t = ld_const(tiles, blockIdx.x);
    if ((int)threadIdx.x >= t.count) return;
    const DevBlock B = ld_const(blocks, t.block);  // descriptor in scalar registers
    const AdjBlock O = ld_const(offs, t.block);
    const int e = t.start + threadIdx.x;
    const int k = e / B.nx, j = e - k * B.nx;
    double* out = rb + (long long)phx::adj::RB * phx::adj::npts<NQ>(ph) * phx::adj::cell_index(O, B, k, j);
    if (B.kind == PHX_GEOM_CARTESIAN) phx::adj::recon_cell<true, NQ>(blocks, B, U, k, j, out, ph);
    else phx::adj::recon_cell<false, NQ>(blocks, B, U, k, j, out, ph);
}

template <int NQ>
__global__ void __launch_bounds__(TILE) k_adj_cell(const DevBlock* blocks, const AdjBlock* offs, const AdjTile* tiles,
                                                   const double* U, const double* fadj, double* slot, Phys ph) {
    const AdjTile t = ld_const(tiles, blockIdx.x);
    if ((int)threadIdx.x >= t.count) return;
    const DevBlock B = ld_const(blocks, t.block);  // descriptor in scalar registers
    const AdjBlock O = ld_const(offs, t.block);
    const int e = t.start + threadIdx.x;
    const int k = e / B.nx, j = e - k * B.nx;
    double* out = slot + phx::adj::SLOT * phx::adj::cell_index(O, B, k, j);
    if (B.kind == PHX_GEOM_CARTESIAN) phx::adj::cell_adj<true, NQ>(blocks, B, U, O, fadj, k, j, out, ph);
    else phx::adj::cell_adj<false, NQ>(blocks, B, U, O, fadj, k, j, out, ph);
}

// Midpoint faces: the reconstruction and its VJP with one lane per cell and variable (4*TILE lanes
// per tile of TILE cells): four times the waves of the per-cell kernels on the same tiles, and the
// registers of one variable's chain.
constexpr int VTILE = 4 * TILE;

__global__ void __launch_bounds__(VTILE) k_adj_recon_v(const DevBlock* blocks, const AdjBlock* offs,
                                                       const AdjTile* tiles, const double* U, double* rb, Phys ph) {
    const AdjTile t = ld_const(tiles, blockIdx.x);
    const int c = (int)threadIdx.x >> 2, v = (int)threadIdx.x & 3;
    if (c >= t.count) return;
    const DevBlock B = ld_const(blocks, t.block);  // descriptor in scalar registers
    const AdjBlock O = ld_const(offs, t.block);
    const int e = t.start + c;
    const int k = e / B.nx, j = e - k * B.nx;
    double* out = rb + (long long)phx::adj::RB * phx::adj::cell_index(O, B, k, j);
    if (B.kind == PHX_GEOM_CARTESIAN) phx::adj::recon_cell_var<true>(blocks, B, U, k, j, v, out, ph);
    else phx::adj::recon_cell_var<false>(blocks, B, U, k, j, v, out, ph);
}

__global__ void __launch_bounds__(VTILE) k_adj_cell_v(const DevBlock* blocks, const AdjBlock* offs,
                                                      const AdjTile* tiles, const double* U, const double* fadj,
                                                      double* slot, Phys ph) {
    const AdjTile t = ld_const(tiles, blockIdx.x);
    const int c = (int)threadIdx.x >> 2, v = (int)threadIdx.x & 3;
    if (c >= t.count) return;
    const DevBlock B = ld_const(blocks, t.block);  // descriptor in scalar registers
    const AdjBlock O = ld_const(offs, t.block);
    const int e = t.start + c;
    const int k = e / B.nx, j = e - k * B.nx;
    double* out = slot + phx::adj::SLOT * phx::adj::cell_index(O, B, k, j);
    if (B.kind == PHX_GEOM_CARTESIAN) phx::adj::cell_adj_var<true>(blocks, B, U, O, fadj, k, j, v, out, ph);
    else phx::adj::cell_adj_var<false>(blocks, B, U, O, fadj, k, j, v, out, ph);
}

// lamU = (dR/dU)^T kappa + extra terms (pointwise; extra may include lamU itself)
template <int NQ>
__global__ void __launch_bounds__(TILE) k_adj_gather(const DevBlock* blocks, const AdjBlock* offs, const AdjTile* tiles,
                                                     const double* slot, const double* fadj, const double* U,
                                                     Phys ph, Terms extra, double* lamU) {
    const AdjTile t = ld_const(tiles, blockIdx.x);
    if ((int)threadIdx.x >= t.count) return;
    const DevBlock B = ld_const(blocks, t.block);  // descriptor in scalar registers
    const int e = t.start + threadIdx.x;
    const int k = e / B.nx, j = e - k * B.nx;
    Q4 lam = (B.kind == PHX_GEOM_CARTESIAN)
                 ? phx::adj::gather_adj<true, NQ>(blocks, offs, t.block, k, j, slot, fadj, U, ph)
                 : phx::adj::gather_adj<false, NQ>(blocks, offs, t.block, k, j, slot, fadj, U, ph);
    if (extra.n > 0) phx::adj::acc(lam, eval_terms(extra, B, B.off / 4 + (long long)k * B.nx + j));
    phx::adj::st(lamU + B.off + 4LL * ((long long)k * B.nx + j), lam);
}

// ---- tangent-linear products (forward mode): three element-wise passes on the same tiles ----
//   k_tan_recon   face states of every cell and their derivatives (16 + 16 doubles)
//   k_tan_face    derivative of every face flux (4 doubles)
//   k_tan_cell    derivative of every cell's residual (the residual is linear in the fluxes)
__global__ void __launch_bounds__(TILE) k_tan_recon(const DevBlock* blocks, const AdjBlock* offs, const AdjTile* tiles,
                                                    const double* U, const double* dU, double* rb, double* tb,
                                                    Phys ph) {
    const AdjTile t = ld_const(tiles, blockIdx.x);
    if ((int)threadIdx.x >= t.count) return;
    const DevBlock B = ld_const(blocks, t.block);  // descriptor in scalar registers
    const AdjBlock O = ld_const(offs, t.block);
    const int e = t.start + threadIdx.x;
    const int k = e / B.nx, j = e - k * B.nx;
    const long long c = phx::adj::RB * phx::adj::cell_index(O, B, k, j);
    if (B.kind == PHX_GEOM_CARTESIAN) phx::tng::recon_cell_d<true>(blocks, B, U, dU, k, j, rb + c, tb + c, ph);
    else phx::tng::recon_cell_d<false>(blocks, B, U, dU, k, j, rb + c, tb + c, ph);
}

template <int FLUX>
__global__ void __launch_bounds__(TILE) k_tan_face(const DevBlock* blocks, const AdjBlock* offs, const AdjTile* tiles,
                                                   const double* U, const double* dU, const double* rb,
                                                   const double* tb, double* dF, Phys ph) {
    const AdjTile t = ld_const(tiles, blockIdx.x);
    if ((int)threadIdx.x >= t.count) return;
    const DevBlock B = ld_const(blocks, t.block);  // descriptor in scalar registers
    const AdjBlock O = ld_const(offs, t.block);
    const long long f = (long long)t.start + threadIdx.x;
    const Q4 d = (B.kind == PHX_GEOM_CARTESIAN)
                     ? phx::tng::face_flux_d<true, FLUX>(blocks, B, U, dU, rb, tb, O, f, ph)
                     : phx::tng::face_flux_d<false, FLUX>(blocks, B, U, dU, rb, tb, O, f, ph);
    phx::adj::st(dF + 4 * (O.face + f), d);
}

__global__ void __launch_bounds__(TILE) k_tan_cell(const DevBlock* blocks, const AdjBlock* offs, const AdjTile* tiles,
                                                   const double* dF, double* out) {
    const AdjTile t = ld_const(tiles, blockIdx.x);
    if ((int)threadIdx.x >= t.count) return;
    const DevBlock B = ld_const(blocks, t.block);  // descriptor in scalar registers
    const AdjBlock O = ld_const(offs, t.block);
    const int e = t.start + threadIdx.x;
    const int k = e / B.nx, j = e - k * B.nx;
    const Q4 r = (B.kind == PHX_GEOM_CARTESIAN) ? phx::adj::cell_residual<true>(B, O, dF, k, j)
                                                : phx::adj::cell_residual<false>(B, O, dF, k, j);
    phx::adj::st(out + B.off + 4LL * ((long long)k * B.nx + j), r);
}

// out = sum of terms at every cell (the stage combinations of the tangent step)
__global__ void __launch_bounds__(TILE) k_tan_combine(const DevBlock* blocks, const AdjTile* tiles, Terms terms,
                                                      double* out) {
    const AdjTile t = ld_const(tiles, blockIdx.x);
    if ((int)threadIdx.x >= t.count) return;
    const DevBlock B = ld_const(blocks, t.block);  // descriptor in scalar registers
    const long long cell = B.off / 4 + t.start + threadIdx.x;
    phx::adj::st(out + 4 * cell, eval_terms(terms, B, cell));
}

template <class T>
int upload(const std::vector<T>& h, T** d) {
    if (*d) { (void)hipFree(*d); *d = nullptr; }
    if (h.empty()) return PHX_OK;
    PHX_CK(hipMalloc(d, h.size() * sizeof(T)));
    PHX_CK(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return PHX_OK;
}

const int OPP[4] = {PHX_W, PHX_E, PHX_S, PHX_N};

// Per-block offsets and launch tiles of the adjoint; requires neighbour links to be mutual.  Run
// once, when the plan is created.
int adj_meta(phx_plan* P) {
    for (int b = 0; b < P->nblocks; ++b) {
        const phx_block_desc& d = P->desc[b];
        for (int s = 0; s < 4; ++s) {
            if (d.side_type[s] != PHX_SIDE_NEIGHBOR) continue;
            const phx_block_desc& o = P->desc[d.neighbor[s]];
            if (o.side_type[OPP[s]] != PHX_SIDE_NEIGHBOR || o.neighbor[OPP[s]] != b) {
                P->adj_meta_error = "adjoint: block " + std::to_string(b) + " has a neighbour link that is not mutual";
                return PHX_OK;  // the forward path does not need mutual links; the adjoint calls refuse the plan
            }
        }
    }
    std::vector<AdjBlock> offs(P->nblocks);
    long long faces = 0;
    P->adj_cell_tiles.clear();
    P->adj_face_tiles.clear();
    for (int b = 0; b < P->nblocks; ++b) {
        const phx_block_desc& d = P->desc[b];
        const long long nc = (long long)d.nx * d.ny;
        const long long nf = (long long)d.ny * (d.nx + 1) + (long long)(d.ny + 1) * d.nx;
        offs[b].cell = d.state_offset / 4;
        offs[b].face = faces;
        faces += nf;
        for (long long e = 0; e < nc; e += TILE)
            P->adj_cell_tiles.push_back({b, (int)e, (int)std::min<long long>(TILE, nc - e)});

        for (long long e = 0; e < nf; e += TILE)
            P->adj_face_tiles.push_back({b, (int)e, (int)std::min<long long>(TILE, nf - e)});
    }
    P->adj_cells = P->state_len / 4;
    P->adj_faces = faces;
    int rc;
    if ((rc = upload(offs, &P->dadj))) return rc;
    if ((rc = upload(P->adj_cell_tiles, &P->dadj_cell_tiles))) return rc;
    if ((rc = upload(P->adj_face_tiles, &P->dadj_face_tiles))) return rc;
    P->adj_meta_ready = true;
    return PHX_OK;
}

// Work arrays of one residual VJP: limited face states per cell, face adjoints, cell slots.
struct AdjBufs {
    double *rb, *face, *slot;
};

// Doubles of each work array of a plan (cells x RB x nq, faces x FA x nq, cells x SLOT).
void adj_sizes(const phx_plan* P, size_t n[3]) {
    const size_t nq = (size_t)std::max(P->ph.nq, 1);
    n[0] = (size_t)P->adj_cells * phx::adj::RB * nq;
    n[1] = (size_t)P->adj_faces * phx::adj::FA * nq;
    n[2] = (size_t)P->adj_cells * phx::adj::SLOT;
}

// The plan's own work arrays, allocated on the first adjoint call.
int adj_setup(phx_plan* P) {
    if (P->adj_ready) return PHX_OK;
    if (!P->adj_meta_ready) return fail(PHX_E_INVALID, P->adj_meta_error);
    size_t n[3];
    adj_sizes(P, n);
    const size_t L = (size_t)P->state_len;
    PHX_CK(hipMalloc(&P->adj_rb, n[0] * sizeof(double)));
    PHX_CK(hipMalloc(&P->adj_face, n[1] * sizeof(double)));
    PHX_CK(hipMalloc(&P->adj_slot, n[2] * sizeof(double)));
    for (int s = 1; s < P->nstages; ++s) PHX_CK(hipMalloc(&P->adj_lam[s], L * sizeof(double)));
    P->adj_ready = true;
    return PHX_OK;
}

// The tangent's stage derivatives K_s (one state array per stage) and the derivative of the current
// stage input, allocated on the first tangent call (besides the adjoint work arrays it shares).
int tan_setup(phx_plan* P) {
    int rc = adj_setup(P);
    if (rc) return rc;
    if (P->tan_u) return PHX_OK;
    const size_t bytes = (size_t)P->state_len * sizeof(double);
    for (int s = 0; s < P->nstages; ++s) PHX_CK(hipMalloc(&P->tan_k[s], bytes));
    PHX_CK(hipMalloc(&P->tan_u, bytes));
    return PHX_OK;
}

void launch_residual_jvp(phx_plan* P, const double* U, const double* dU, double* out, hipStream_t st) {
    const dim3 blk(TILE);
    const dim3 gc((unsigned)P->adj_cell_tiles.size()), gf((unsigned)P->adj_face_tiles.size());
    // face states in adj_rb, their derivatives in adj_slot (RB of its SLOT doubles per cell), face
    // flux derivatives in adj_face (4 of its FA doubles per face)
    hipLaunchKernelGGL(k_tan_recon, gc, blk, 0, st, P->dblocks, P->dadj, P->dadj_cell_tiles, U, dU, P->adj_rb,
                       P->adj_slot, P->ph);
    if (P->flux == PHX_FLUX_ROE)
        hipLaunchKernelGGL(k_tan_face<0>, gf, blk, 0, st, P->dblocks, P->dadj, P->dadj_face_tiles, U, dU, P->adj_rb,
                           P->adj_slot, P->adj_face, P->ph);
    else
        hipLaunchKernelGGL(k_tan_face<1>, gf, blk, 0, st, P->dblocks, P->dadj, P->dadj_face_tiles, U, dU, P->adj_rb,
                           P->adj_slot, P->adj_face, P->ph);
    hipLaunchKernelGGL(k_tan_cell, gc, blk, 0, st, P->dblocks, P->dadj, P->dadj_cell_tiles, P->adj_face, out);
}

int tan_check(phx_plan* P) {
    if (P->ph.nq > 1) return fail(PHX_E_INVALID, "tangent-linear products support midpoint faces only (quadrature_points 1)");
    return tan_setup(P);
}

// ---- static work array of phx_residual_vjp_noalloc ----
// One array per GPU: every device that runs the library holds its own copy until the process
// ends.  In host code `scratch` names a host-side stand-in, never the device memory: the device
// address comes from hipGetSymbolAddress, taken once per device at plan creation (with the plan's
// device current) and kept in scratch_ptr.
constexpr size_t NOALLOC_DOUBLES = PHX_NOALLOC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double* scratch_ptr[MAX_DEVICES] = {};

int scratch_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) return fail(PHX_E_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (scratch_ptr[device]) return PHX_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(PHX_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    scratch_ptr[device] = static_cast<double*>(p);
    return PHX_OK;
}

double* scratch_of(int device) {
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(scratch_mu);
    return scratch_ptr[device];
}

template <int NQ>
void launch_residual_vjp_nq(phx_plan* P, const double* U, const Terms& kappa, const Terms& extra, double* lamU,
                            const AdjBufs& w, hipStream_t st) {
    const dim3 blk(TILE), vblk(VTILE);
    const dim3 gc((unsigned)P->adj_cell_tiles.size()), gf((unsigned)P->adj_face_tiles.size());
    if (NQ == 1)
        hipLaunchKernelGGL(k_adj_recon_v, gc, vblk, 0, st, P->dblocks, P->dadj, P->dadj_cell_tiles, U, w.rb, P->ph);
    else
        hipLaunchKernelGGL(k_adj_recon<NQ>, gc, blk, 0, st, P->dblocks, P->dadj, P->dadj_cell_tiles, U, w.rb,
                           P->ph);
    if (P->flux == PHX_FLUX_ROE)
        hipLaunchKernelGGL((k_adj_face<0, NQ>), gf, blk, 0, st, P->dblocks, P->dadj, P->dadj_face_tiles, U, w.rb,
                           kappa, w.face, P->ph);
    else
        hipLaunchKernelGGL((k_adj_face<1, NQ>), gf, blk, 0, st, P->dblocks, P->dadj, P->dadj_face_tiles, U, w.rb,
                           kappa, w.face, P->ph);
    if (NQ == 1)
        hipLaunchKernelGGL(k_adj_cell_v, gc, vblk, 0, st, P->dblocks, P->dadj, P->dadj_cell_tiles, U, w.face,
                           w.slot, P->ph);
    else
        hipLaunchKernelGGL(k_adj_cell<NQ>, gc, blk, 0, st, P->dblocks, P->dadj, P->dadj_cell_tiles, U, w.face,
                           w.slot, P->ph);
    hipLaunchKernelGGL(k_adj_gather<NQ>, gc, blk, 0, st, P->dblocks, P->dadj, P->dadj_cell_tiles, w.slot,
                       w.face, U, P->ph, extra, lamU);
}

void launch_residual_vjp(phx_plan* P, const double* U, const Terms& kappa, const Terms& extra, double* lamU,
                         const AdjBufs& w, hipStream_t st) {
    if (P->ph.nq > 1) launch_residual_vjp_nq<phx::QMAX>(P, U, kappa, extra, lamU, w, st);
    else launch_residual_vjp_nq<1>(P, U, kappa, extra, lamU, w, st);
}


}  // namespace

void phx::host::adjoint_release(phx_plan* P) {
    void* ptrs[] = {P->dadj, P->dadj_cell_tiles, P->dadj_face_tiles, P->adj_rb, P->adj_face, P->adj_slot};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (int s = 0; s < PHX_MAX_STAGES; ++s)
        if (P->adj_lam[s]) (void)hipFree(P->adj_lam[s]);
    P->dadj = nullptr;
    P->dadj_cell_tiles = P->dadj_face_tiles = nullptr;
    P->adj_rb = P->adj_face = P->adj_slot = nullptr;
    for (int s = 0; s < PHX_MAX_STAGES; ++s) P->adj_lam[s] = nullptr;
    for (int s = 0; s < PHX_MAX_STAGES; ++s) {
        if (P->tan_k[s]) (void)hipFree(P->tan_k[s]);
        P->tan_k[s] = nullptr;
    }
    if (P->tan_u) (void)hipFree(P->tan_u);
    P->tan_u = nullptr;
    P->adj_ready = false;
    P->adj_meta_ready = false;
}

int phx::host::adjoint_prepare(phx_plan* P) {
    const int rc = adj_meta(P);
    if (rc) return rc;
    return scratch_setup(P->device);
}

extern "C" {

int phx_residual_vjp(phx_plan* P, const double* U, const double* lamR, double* lamU, void* stream) {
    if (!P || !U || !lamR || !lamU) return fail(PHX_E_INVALID, "null argument");
    int rc = adj_setup(P);
    if (rc) return rc;
    Terms seed, none;
    memset(&seed, 0, sizeof(seed));
    memset(&none, 0, sizeof(none));
    seed.in[0] = lamR;
    seed.c[0] = 1.0;
    seed.n = 1;
    launch_residual_vjp(P, U, seed, none, lamU, AdjBufs{P->adj_rb, P->adj_face, P->adj_slot}, (hipStream_t)stream);
    PHX_CK(hipGetLastError());
    return PHX_OK;
}

// phx_residual_vjp with its work arrays in the static array `scratch` (no allocation at all): the
// same kernels on the same launch tiles, so the result is bit-identical.  Every array element the
// kernels read is written earlier in the same call, so nothing has to be cleared.
int phx_residual_vjp_noalloc(phx_plan* P, const double* U, const double* lamR, double* lamU, void* stream) {
    if (!P || !U || !lamR || !lamU) return fail(PHX_E_INVALID, "null argument");
    if (!P->adj_meta_ready) return fail(PHX_E_INVALID, P->adj_meta_error);
    int dev = -1;
    PHX_CK(hipGetDevice(&dev));
    if (dev != P->device)
        return fail(PHX_E_INVALID, "phx_residual_vjp_noalloc: the current device is not the plan's device");
    double* base = scratch_of(P->device);
    if (!base) return fail(PHX_E_STATE, "phx_residual_vjp_noalloc: static work array not set up on this device");
    size_t n[3];
    adj_sizes(P, n);
    const size_t align = 256 / sizeof(double);
    size_t off[3], total = 0;
    for (int i = 0; i < 3; ++i) {
        off[i] = total;
        total += (n[i] + align - 1) / align * align;
    }
    if (total > NOALLOC_DOUBLES)
        return fail(PHX_E_INVALID, "phx_residual_vjp_noalloc: the plan needs " + std::to_string(total * sizeof(double)) +
                                       " bytes of adjoint work arrays, more than the static array's " +
                                       std::to_string((size_t)PHX_NOALLOC_SCRATCH_BYTES) +
                                       " (use phx_residual_vjp)");
    Terms seed, none;
    memset(&seed, 0, sizeof(seed));
    memset(&none, 0, sizeof(none));
    seed.in[0] = lamR;
    seed.c[0] = 1.0;
    seed.n = 1;
    launch_residual_vjp(P, U, seed, none, lamU, AdjBufs{base + off[0], base + off[1], base + off[2]},
                        (hipStream_t)stream);
    PHX_CK(hipGetLastError());
    return PHX_OK;
}

// One explicit RK step U_S = U_0 + dt sum_m a[S-1][m] K_m, U_i = U_0 + dt sum_m a[i-1][m] K_m,
// K_m = R(U_m), with dt frozen.  Reverse: kappa_i = dt sum_{l>i} a[l-1][i] lambda_l,
// lambda_i = (dR/dU(U_i))^T kappa_i for i >= 1, and lambda_0 = lambda_S + sum_{l=1}^{S-1} lambda_l
// + (dR/dU(U_0))^T kappa_0.
int phx_step_vjp(phx_plan* P, const double* stages, const double* dt_dev, double* lam, void* stream) {
    if (!P || !stages || !dt_dev || !lam) return fail(PHX_E_INVALID, "null argument");
    int rc = adj_setup(P);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const int S = P->nstages;
    const size_t L = (size_t)P->state_len;
    const double* a = P->butcher.data();
    auto lam_of = [&](int l) -> const double* { return l == S ? lam : P->adj_lam[l]; };
    const AdjBufs W{P->adj_rb, P->adj_face, P->adj_slot};
    for (int i = S - 1; i >= 0; --i) {
        // seed kappa_i, evaluated inside the face kernel
        Terms kap, extra;
        memset(&kap, 0, sizeof(kap));
        memset(&extra, 0, sizeof(extra));
        for (int l = i + 1; l <= S; ++l) {
            const double c = a[(size_t)(l - 1) * S + i];
            if (c == 0.0) continue;
            kap.in[kap.n] = lam_of(l);
            kap.c[kap.n] = c;
            kap.scaled[kap.n] = 1;
            ++kap.n;
        }
        kap.dt = dt_dev;
        if (i > 0) {
            launch_residual_vjp(P, stages + (size_t)i * L, kap, extra, P->adj_lam[i], W, st);
        } else {
            // lambda_0 = lambda_S + sum_{l=1}^{S-1} lambda_l + (dR/dU(U_0))^T kappa_0, written over lam
            extra.in[extra.n] = lam;
            extra.c[extra.n++] = 1.0;
            for (int l = 1; l < S; ++l) {
                extra.in[extra.n] = P->adj_lam[l];
                extra.c[extra.n++] = 1.0;
            }
            launch_residual_vjp(P, stages, kap, extra, lam, W, st);
        }
    }
    PHX_CK(hipGetLastError());
    return PHX_OK;
}

// Forward-mode product of the residual: dR = (dR/dU) dU at U.
int phx_residual_jvp(phx_plan* P, const double* U, const double* dU, double* dR, void* stream) {
    if (!P || !U || !dU || !dR) return fail(PHX_E_INVALID, "null argument");
    int rc = tan_check(P);
    if (rc) return rc;
    launch_residual_jvp(P, U, dU, dR, (hipStream_t)stream);
    PHX_CK(hipGetLastError());
    return PHX_OK;
}

// Forward-mode product of one RK step with a frozen time step: dU_i = dU_0 + dt sum_m a[i-1][m] dK_m,
// dK_m = (dR/dU)(U_m) dU_m, and dU_S (the step's output) written over dU.
int phx_step_jvp(phx_plan* P, const double* stages, const double* dt_dev, double* dU, void* stream) {
    if (!P || !stages || !dt_dev || !dU) return fail(PHX_E_INVALID, "null argument");
    int rc = tan_check(P);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const int S = P->nstages;
    const size_t L = (size_t)P->state_len;
    const double* a = P->butcher.data();
    const dim3 gc((unsigned)P->adj_cell_tiles.size());
    for (int s = 0; s < S; ++s) {
        launch_residual_jvp(P, stages + (size_t)s * L, s == 0 ? dU : P->tan_u, P->tan_k[s], st);
        Terms T;
        memset(&T, 0, sizeof(T));
        T.in[T.n] = dU;
        T.c[T.n++] = 1.0;
        for (int m = 0; m <= s; ++m) {
            const double c = a[(size_t)s * S + m];
            if (c == 0.0) continue;
            T.in[T.n] = P->tan_k[m];
            T.c[T.n] = c;
            T.scaled[T.n++] = 1;
        }
        T.dt = dt_dev;
        hipLaunchKernelGGL(k_tan_combine, gc, dim3(TILE), 0, st, P->dblocks, P->dadj_cell_tiles, T,
                           s + 1 < S ? P->tan_u : dU);
    }
    PHX_CK(hipGetLastError());
    return PHX_OK;
}

}  // extern "C"
