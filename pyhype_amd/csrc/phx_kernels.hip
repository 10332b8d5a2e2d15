// phx_kernels.hip -- gfx950 kernels of the pyHype MUSCL/RK backend and the C ABI declared in
// include/pyhype_hip.h.
//
// Stage kernel design ("rolling rows"): one wavefront owns a strip of 60 columns of one block
// plus two halo lanes on each side and walks up a range of rows.  Per row it loads one row of the
// stage state, reconstructs each cell once (Green-Gauss + Venkatakrishnan + limited face states),
// evaluates each x-face flux once (lane j computes face j-1|j and hands it to lane j-1 through a
// cross-lane shuffle), evaluates the y-face between this row and the previous one in registers
// (same lane), finishes the residual of the previous row and applies every Butcher update of the
// stage in the same pass.  Ghost cells (inter-block neighbours, reflected walls, outlets) are
// produced on the fly from the stage state, so there is no separate ghost-fill pass.  The last
// stage of a step also reduces the CFL time step of the next step and the realizability flag.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/pyhype_hip.h"
#include "phx_adjoint_ops.hpp"
#include "phx_device.hpp"
#include "phx_numerics.hpp"
#include "phx_plan.hpp"

using phx::Phys;
using phx::Q4;
using namespace phx::dev;
using namespace phx::host;

namespace {

constexpr int WAVE = 64;
constexpr int HALO = 2;
constexpr int STRIP = WAVE - 2 * HALO;  // owned columns per wavefront
constexpr unsigned long long DT_INF_BITS = 0x7FF0000000000000ull;  // +inf
constexpr int GRAPH_STEPS = 8;  // steps captured in the multi-step graph of phx_run
// The CFL minimum of a case is accumulated in nslot partial minima (a power of two, at most
// DT_SLOTS), each on a cache line of its own (DT_STRIDE words apart): the waves of a stage spread
// their atomicMin over the slots instead of serialising on one address, and k_finalize takes the
// minimum over the slots (an exact operation, independent of the order).  Plans with many cases
// use fewer slots per case (their waves are spread over the cases already).
constexpr int DT_SLOTS = 64;
constexpr int DT_STRIDE = 16;
__host__ __device__ constexpr long dt_slot(int c, int slot, int nslot) { return ((long)c * nslot + slot) * DT_STRIDE; }
__device__ __forceinline__ int wave_slot(int nslot) {
    return (int)((blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE) & (nslot - 1));
}
int dt_slots_for(int ncases) {
    int n = DT_SLOTS;
    while (n > 4 && (long)n * ncases > 256) n >>= 1;
    return n;
}

// Every per-item function below is __forceinline__, so that no kernel makes a device-function call.
// Outlined (s_swappc) item functions of the HLLL Gauss-face and generic shared-halo kernels, which
// appeared when the HLLL norm grew, returned NaN residuals on gfx950 (ROCm 7.2) where the same
// source inlined, and a stand-alone kernel of the same flux, are bit-exact.
struct StageArgs {
    const DevBlock* blocks;
    const WorkItem* items;
    int nitems;
    int nout;
    int state_out;
    const double* Uin;
    const double* dt;
    double* Rout;
    double* dst[MAX_OUT];
    const double* src[MAX_OUT];
    double a[MAX_OUT];
    unsigned long long* dtmin;
    int* bad;          // per case: first step whose result was unrealizable (INT_MAX = none)
    const int* step;   // index of the step being computed (1-based)
    int skip_dead;     // skip the blocks of cases that failed at an earlier step
    int rows_wave;     // shared-halo stage: rows per wave
    double* sn;        // split stage: limited N and S states per cell (8 doubles, offset 2*off)
    double* p1;        // split stage: LW*(2 FW) - LE*(2 FE) per cell (4 doubles, state layout)
    double* yf;        // split stage: y-face fluxes (DevBlock::yoff)
    double* rb;        // element-wise stage: limited face states E, W, N, S per cell (16 doubles)
    double* ff;        // element-wise stage: face fluxes (4 doubles per face, DevBlock::foff)
    int nslot;         // partial CFL minima per case
    int fq;            // flat y-face stage: LDS faces per lane
    Phys ph;
};

// Element i of an array that no kernel writes, as a reference through the constant address
// space: uniform field reads become scalar loads that the compiler may repeat instead of keeping
// the whole element in registers.
template <class T>
__device__ __forceinline__ const T& const_ref(const T* p, long i) {
    return *(const T*)(((const __attribute__((address_space(4))) T*)p) + i);
}

// cross-lane moves of a 4-vector
__device__ __forceinline__ Q4 shfl_up1(const Q4& q) {
    Q4 r;
#pragma unroll
    for (int k = 0; k < 4; ++k) r.v[k] = __shfl_up(q.v[k], 1, WAVE);
    return r;
}
__device__ __forceinline__ Q4 shfl_down1(const Q4& q) {
    Q4 r;
#pragma unroll
    for (int k = 0; k < 4; ++k) r.v[k] = __shfl_down(q.v[k], 1, WAVE);
    return r;
}

// Outputs of the stage program are written in chunks of EMIT_CHUNK: the chunk's sources are loaded
// together before its stores (an output's source is never a buffer an earlier output of the stage
// writes), and the chunk loop is not unrolled, so the outputs' pointers and coefficients are read
// from the kernel arguments per chunk instead of occupying scalar registers for all MAX_OUT slots.
constexpr int EMIT_CHUNK = 3;

// Sources of the first output chunk of owned cell (kk, j), loaded ahead of the row's arithmetic so
// that their latency overlaps it.  Only this lane writes these cells in this stage (after it has
// read them), so the early read returns what emit_row would read.
__device__ __forceinline__ void prefetch_sources(const StageArgs& A, const DevBlock& B, int kk, int j,
                                                 Q4 (&pre)[EMIT_CHUNK]) {
    const long idx = B.off + 4 * ((long)kk * B.nx + j);
#pragma unroll
    for (int q = 0; q < EMIT_CHUNK; ++q)
        if (q < A.nout) pre[q] = ld4(A.src[q] + idx);
}

// Residual of owned cell (kk, j) -> every Butcher output of the stage program
// (explicit_runge_kutta.py:63-80); on the last stage also the cell's CFL quantity and realizability.
// PF: the first chunk's sources were loaded by prefetch_sources into pre; on a Cartesian block dxc
// is the cell's column width and the row height a scalar load.
template <bool CART, bool PF = false>
__device__ __forceinline__ void emit_row(const StageArgs& A, const DevBlock& B, int kk, int j, double dt, const Q4& Res,
                                         double& lane_dt, bool& lane_bad, const Q4* pre = nullptr,
                                         double dxc = 0.0) {
    const long idx = B.off + 4 * ((long)kk * B.nx + j);
    if (A.Rout) st4(A.Rout + idx, Res);
#pragma unroll 1
    for (int m0 = 0; m0 < A.nout; m0 += EMIT_CHUNK) {
        Q4 s[EMIT_CHUNK];
        if (PF && m0 == 0) {
#pragma unroll
            for (int q = 0; q < EMIT_CHUNK; ++q) s[q] = pre[q];
        } else {
#pragma unroll
            for (int q = 0; q < EMIT_CHUNK; ++q)
                if (m0 + q < A.nout) s[q] = ld4(A.src[m0 + q] + idx);
        }
#pragma unroll
        for (int q = 0; q < EMIT_CHUNK; ++q) {
            const int m = m0 + q;
            if (m >= A.nout) break;
            Q4 o;
            const double am = A.a[m];
            if (am == 0.0) {
                o = s[q];
            } else {
                const double cdt = dt * am;  // dt * a[stage][step] (explicit_runge_kutta.py:72)
#pragma unroll
                for (int c = 0; c < 4; ++c) o.v[c] = s[q].v[c] + cdt * Res.v[c];
            }
            st4(A.dst[m] + idx, o);
            if (m == A.state_out && A.dtmin != nullptr) {
                double dx, dy;
                if (PF && CART) {
                    dx = dxc;
                    dy = CartGT<true>{B.geom, B.nx, B.ny}.row(R_dy, kk);
                } else {
                    cell_dxdy<CART>(B, kk, j, dx, dy);
                }
                const double t = phx::cell_dt(o, dx, dy, A.ph);
                if (!(o.v[0] > 0.0) || !(o.v[3] > 0.0) || !(t == t)) lane_bad = true;
                else lane_dt = fmin(lane_dt, t);
            }
        }
    }
}

// wave-level min of the CFL quantities and any-unrealizable of one work item -> per-case atomics
__device__ __forceinline__ void reduce_item(const StageArgs& A, const DevBlock& B, int lane, double lane_dt,
                                            bool lane_bad) {
    if (A.dtmin != nullptr) {
        double m = lane_dt;
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) m = fmin(m, __shfl_xor(m, o, WAVE));
        const bool anybad = __any(lane_bad);
        if (lane == 0) {
            const unsigned long long bits = (unsigned long long)__double_as_longlong(m);
            if (bits < DT_INF_BITS) atomicMin(&A.dtmin[dt_slot(B.case_id, wave_slot(A.nslot), A.nslot)], bits);
            if (anybad) atomicMin(&A.bad[B.case_id], *A.step);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// one work item of one RK stage
// ---------------------------------------------------------------------------------------------
// PF (the instances with a budget of 2 waves per SIMD, which have registers to spare): the block
// descriptor is read through the constant address space (scalar loads), and each row iteration
// issues its loads in the order they are needed, before its arithmetic: the row's geometry and face
// lengths, then the next row's states and the sources of the row it emits.  Vector-memory waits
// are in issue order, so the wait for the geometry leaves the later loads in flight.  On Cartesian
// blocks the column arrays are loaded once per item and the cell areas one row ahead, so the only
// geometry a row iteration waits for was requested an iteration earlier.
template <bool CART, int FLUX, bool GEN, bool PF = false>
__device__ __forceinline__ void stage_item(const StageArgs& A, const WorkItem it) {
    const DevBlock& B = PF ? const_ref(A.blocks, it.block) : A.blocks[it.block];
    // a case that became unrealizable is frozen: the reference aborts that run (Euler2D.py:144-152)
    if (A.skip_dead && A.bad[B.case_id] < *A.step) return;
    const int lane = threadIdx.x & (WAVE - 1);
    const int nx = B.nx, ny = B.ny;
    const int j = it.j0 - HALO + lane;
    const bool valid = (j >= -1) && (j <= nx);
    const bool inner = (j >= 0) && (j < nx);
    const bool gcol = (j == -1) || (j == nx);
    const bool recon = inner && (j >= it.j0 - 1) && (j <= it.j0 + STRIP);
    const bool xface = (j >= it.j0) && (j <= min(it.j0 + STRIP, nx));
    const bool owner = inner && (j >= it.j0) && (j < it.j0 + STRIP);
    const double dt = (A.dt != nullptr) ? ld_const(A.dt, B.case_id) : 0.0;
    const Phys ph = A.ph;

    const int kr0 = max(it.r0 - 1, 0);
    const int kr1 = min(it.r1, ny);  // row ny is a virtual row closing the north boundary face

    Q4 Um = phx::q4zero(), Uc = phx::q4zero(), Up = phx::q4zero();
    if (inner) Um = fetch_r<CART, GEN>(A.blocks, B, A.Uin, kr0 - 1, j, ph);
    if (valid) Uc = fetch_r<CART, GEN>(A.blocks, B, A.Uin, kr0, j, ph);

    Q4 QNprev = phx::q4zero(), part = phx::q4zero();
    double prevLN = 0.0, prevA = 1.0;
    double lane_dt = __longlong_as_double((long long)DT_INF_BITS);
    bool lane_bad = false;

    // Cartesian blocks: the column arrays once per item, the cell areas one row ahead, the row
    // arrays as scalar loads
    const CartGT<true> CG{B.geom, nx, ny};
    CartCols cols{};
    double Acur = 1.0, Anext = 1.0;
    if (PF && CART && inner) {
        cols = cart_cols<true>(B, j);
        Acur = CG.A(kr0, j);
    }

    Q4 pre[EMIT_CHUNK];
    for (int k = kr0; k <= kr1; ++k) {
        const bool real = (k < ny);
        const bool xrow = (k >= it.r0) && (k < it.r1);
        phx::CellGeo g;
        double LW = 0.0, LE = 0.0, LS = 0.0, LN = 0.0, Ac = 1.0;
        if (PF && CART) {
            if (real && recon) g = cart_cell_geo<true>(B, cols, k, Acur);
            if (xrow && owner) {
                LW = CG.row(R_LW, k); LE = CG.row(R_LE, k);
                LS = cols.LS; LN = cols.LN;
                Ac = Acur;
            }
        } else if (PF) {
            if (real && recon) g = cell_geo<CART>(B, k, j);
            if (xrow && owner) cell_L<CART>(B, k, j, LW, LE, LS, LN, Ac);
        }
        if (real) {
            if (inner) Up = fetch_r<CART, GEN>(A.blocks, B, A.Uin, k + 1, j, ph);
            else if (gcol && k + 1 <= kr1 && k + 1 < ny) Up = fetch_r<CART, GEN>(A.blocks, B, A.Uin, k + 1, j, ph);
        }
        if (PF && owner && k - 1 >= it.r0) prefetch_sources(A, B, k - 1, j, pre);
        if (PF && CART && inner && k + 1 < ny) Anext = CG.A(k + 1, j);
        const Q4 Uw = shfl_up1(Uc);
        const Q4 Ue = shfl_down1(Uc);
        phx::Recon r;
        r.E = r.W = r.N = r.S = phx::q4zero();
        if (real && recon) {
            if (!PF) g = cell_geo<CART>(B, k, j);
            r = phx::reconstruct<GEN>(Uc, Uw, Ue, Um, Up, g, ph.limiter);
        }
        const Q4 qsend = (j == -1) ? Uc : r.E;
        const Q4 QEl = shfl_up1(qsend);  // limited east state of cell j-1 (ghost value for j == 0)
        Q4 FW = phx::q4zero();
        if (xrow && xface) {
            const Q4 L = (j == 0) ? outer_state<CART>(B, PHX_W, r.W, QEl, k, j) : QEl;
            const Q4 R = (j == nx) ? outer_state<CART>(B, PHX_E, QEl, Uc, k, j) : r.W;
            FW = xflux<CART, FLUX, GEN>(B, L, R, k, j, ph);
        }
        const Q4 FE = shfl_down1(FW);  // flux of face j+1 (east face of cell j)

        if (k >= it.r0 && owner) {
            // face between rows k-1 and k (south face of row k, north face of row k-1)
            const Q4 L = (k == 0) ? outer_state<CART>(B, PHX_S, r.S, Um, k, j) : QNprev;
            const Q4 R = (k == ny) ? outer_state<CART>(B, PHX_N, QNprev, Uc, k - 1, j) : r.S;
            const Q4 F = yflux<CART, FLUX, GEN>(B, L, R, k, j, ph);
            if (k - 1 >= it.r0) {
                // residual of row k-1: 0.5*(W - E + S - N)/A with I = L*(2*F) (fvm/base.py:141-190)
                const int kk = k - 1;
                Q4 Res;
#pragma unroll
                for (int c = 0; c < 4; ++c) Res.v[c] = (0.5 * (part.v[c] - prevLN * (2.0 * F.v[c]))) / prevA;
                emit_row<CART, PF>(A, B, kk, j, dt, Res, lane_dt, lane_bad, pre, cols.dx);
            }
            if (xrow) {
                if (!PF) cell_L<CART>(B, k, j, LW, LE, LS, LN, Ac);
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    part.v[c] = (LW * (2.0 * FW.v[c]) - LE * (2.0 * FE.v[c])) + LS * (2.0 * F.v[c]);
                prevLN = LN;
                prevA = Ac;
            }
        }
        QNprev = r.N;
        Um = Uc;
        Uc = Up;
        Acur = Anext;
    }

    reduce_item(A, B, lane, lane_dt, lane_bad);
}

// Work distribution: 8 XCDs receive workgroups round-robin; each XCD gets a contiguous range of the
// (spatially ordered) work list so that halo rows/columns shared by neighbouring items are reused
// from the same L2.
__device__ __forceinline__ void item_range(int n, int& lo, int& hi, int& first, int& stride) {
    const int G = gridDim.x, w = blockIdx.x;
    if (G >= 8) {
        const int x = w & 7;
        lo = (int)(((long)n * x) >> 3);
        hi = (int)(((long)n * (x + 1)) >> 3);
        first = lo + (w >> 3);
        stride = (G - x + 7) >> 3;
    } else {
        lo = 0; hi = n; first = w; stride = G;
    }
}

// MINW: minimum waves per SIMD the register allocation must allow (2 or 3; tuned per workload)
// GEN: limiter other than Venkatakrishnan, selected at run time
template <bool CART, int FLUX, int MINW, bool GEN>
__global__ void __launch_bounds__(WAVE, MINW) k_stage(StageArgs A) {
    int lo, hi, first, stride;
    item_range(A.nitems, lo, hi, first, stride);
    for (int i = first; i < hi; i += stride) stage_item<CART, FLUX, GEN, MINW == 2>(A, ld_const(A.items, i));
}

// ---------------------------------------------------------------------------------------------
// shared-halo stage: a workgroup of WGW waves takes WGW * rows_wave consecutive rows of one strip,
// wave w the rows [a, b) = [r0 + w*rows_wave, ...).  Only the workgroup's outermost waves
// reconstruct a halo row; the face between two waves of the workgroup is evaluated once, by the
// upper wave, from the lower wave's last N state passed through LDS, and its flux is passed back.
// Same arithmetic and association order as stage_item (the first and last rows of a wave complete
// after the exchange).
// ---------------------------------------------------------------------------------------------
constexpr int WGW = 4;

template <bool CART, int FLUX, bool GEN>
__device__ __forceinline__ void stage_wg_item(const StageArgs& A, const WorkItem it, Q4* lds_N, Q4* lds_F) {
    const DevBlock& B = A.blocks[it.block];
    if (A.skip_dead && A.bad[B.case_id] < *A.step) return;  // uniform over the workgroup
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = threadIdx.x / WAVE;
    const int nx = B.nx, ny = B.ny;
    const int R = A.rows_wave;
    const int a = it.r0 + w * R, b = min(a + R, it.r1);
    const int wtop = (it.r1 - it.r0 - 1) / R;
    const bool active = a < it.r1;
    const bool bottom = (w == 0), top = (w == wtop);
    const int j = it.j0 - HALO + lane;
    const bool valid = (j >= -1) && (j <= nx);
    const bool inner = (j >= 0) && (j < nx);
    const bool gcol = (j == -1) || (j == nx);
    const bool recon = inner && (j >= it.j0 - 1) && (j <= it.j0 + STRIP);
    const bool xface = (j >= it.j0) && (j <= min(it.j0 + STRIP, nx));
    const bool owner = inner && (j >= it.j0) && (j < it.j0 + STRIP);
    const double dt = (A.dt != nullptr) ? ld_const(A.dt, B.case_id) : 0.0;
    const Phys ph = A.ph;

    double lane_dt = __longlong_as_double((long long)DT_INF_BITS);
    bool lane_bad = false;
    // first row of a wave above the bottom: its S face waits for the exchange
    Q4 first_S = phx::q4zero(), first_part = phx::q4zero(), first_FN = phx::q4zero();
    double first_LS = 0.0, first_LN = 0.0, first_A = 1.0;
    // last row of a wave below the top: its N face waits for the exchange
    Q4 part = phx::q4zero(), QNprev = phx::q4zero();
    double prevLN = 0.0, prevA = 1.0;

    if (active) {
        const int kr0 = bottom ? max(a - 1, 0) : a;
        const int kr1 = top ? min(it.r1, ny) : b - 1;
        Q4 Um = phx::q4zero(), Uc = phx::q4zero(), Up = phx::q4zero();
        if (inner) Um = fetch_r<CART, GEN>(A.blocks, B, A.Uin, kr0 - 1, j, ph);
        if (valid) Uc = fetch_r<CART, GEN>(A.blocks, B, A.Uin, kr0, j, ph);
        for (int k = kr0; k <= kr1; ++k) {
            const bool real = (k < ny);
            if (real) {
                if (inner) Up = fetch_r<CART, GEN>(A.blocks, B, A.Uin, k + 1, j, ph);
                else if (gcol && k + 1 <= kr1 && k + 1 < ny) Up = fetch_r<CART, GEN>(A.blocks, B, A.Uin, k + 1, j, ph);
            }
            const Q4 Uw = shfl_up1(Uc);
            const Q4 Ue = shfl_down1(Uc);
            phx::Recon r;
            r.E = r.W = r.N = r.S = phx::q4zero();
            if (real && recon) {
                const phx::CellGeo g = cell_geo<CART>(B, k, j);
                r = phx::reconstruct<GEN>(Uc, Uw, Ue, Um, Up, g, ph.limiter);
            }
            const Q4 qsend = (j == -1) ? Uc : r.E;
            const Q4 QEl = shfl_up1(qsend);
            const bool xrow = (k >= a) && (k < b);
            Q4 FW = phx::q4zero();
            if (xrow && xface) {
                const Q4 L = (j == 0) ? outer_state<CART>(B, PHX_W, r.W, QEl, k, j) : QEl;
                const Q4 R2 = (j == nx) ? outer_state<CART>(B, PHX_E, QEl, Uc, k, j) : r.W;
                FW = xflux<CART, FLUX, GEN>(B, L, R2, k, j, ph);
            }
            const Q4 FE = shfl_down1(FW);
            if (k >= a && owner) {
                const bool defer_S = (k == a) && !bottom;
                Q4 F = phx::q4zero();
                if (!defer_S) {
                    const Q4 L = (k == 0) ? outer_state<CART>(B, PHX_S, r.S, Um, k, j) : QNprev;
                    const Q4 R2 = (k == ny) ? outer_state<CART>(B, PHX_N, QNprev, Uc, k - 1, j) : r.S;
                    F = yflux<CART, FLUX, GEN>(B, L, R2, k, j, ph);
                    if (k - 1 >= a) {
                        if (k - 1 == a && !bottom) {
                            first_FN = F;
                        } else {
                            Q4 Res;
#pragma unroll
                            for (int c = 0; c < 4; ++c) Res.v[c] = (0.5 * (part.v[c] - prevLN * (2.0 * F.v[c]))) / prevA;
                            emit_row<CART>(A, B, k - 1, j, dt, Res, lane_dt, lane_bad);
                        }
                    }
                }
                if (xrow) {
                    double LW, LE, LS, LN, Ac;
                    cell_L<CART>(B, k, j, LW, LE, LS, LN, Ac);
                    if (defer_S) {
#pragma unroll
                        for (int c = 0; c < 4; ++c) first_part.v[c] = LW * (2.0 * FW.v[c]) - LE * (2.0 * FE.v[c]);
                        first_S = r.S;
                        first_LS = LS;
                        first_LN = LN;
                        first_A = Ac;
                    } else {
#pragma unroll
                        for (int c = 0; c < 4; ++c)
                            part.v[c] = (LW * (2.0 * FW.v[c]) - LE * (2.0 * FE.v[c])) + LS * (2.0 * F.v[c]);
                    }
                    prevLN = LN;
                    prevA = Ac;
                }
            }
            QNprev = r.N;
            Um = Uc;
            Uc = Up;
        }
        if (!top && owner) lds_N[w * WAVE + lane] = QNprev;  // N state of row b-1
    }
    __syncthreads();
    Q4 FS_first = phx::q4zero();
    if (active && !bottom && owner) {
        // face between a-1 (last row of wave w-1) and a
        FS_first = yflux<CART, FLUX, GEN>(B, lds_N[(w - 1) * WAVE + lane], first_S, a, j, ph);
        lds_F[w * WAVE + lane] = FS_first;
    }
    __syncthreads();
    if (active && owner) {
        const bool single = (b - 1 == a);
        if (!bottom) {
            const Q4 FN = (single && !top) ? lds_F[(w + 1) * WAVE + lane] : first_FN;
            Q4 Res;
#pragma unroll
            for (int c = 0; c < 4; ++c)
                Res.v[c] = (0.5 * ((first_part.v[c] + first_LS * (2.0 * FS_first.v[c])) - first_LN * (2.0 * FN.v[c]))) /
                           first_A;
            emit_row<CART>(A, B, a, j, dt, Res, lane_dt, lane_bad);
        }
        if (!top && !(single && !bottom)) {
            const Q4 FN = lds_F[(w + 1) * WAVE + lane];
            Q4 Res;
#pragma unroll
            for (int c = 0; c < 4; ++c) Res.v[c] = (0.5 * (part.v[c] - prevLN * (2.0 * FN.v[c]))) / prevA;
            emit_row<CART>(A, B, b - 1, j, dt, Res, lane_dt, lane_bad);
        }
    }
    reduce_item(A, B, lane, lane_dt, lane_bad);
}

template <bool CART, int FLUX, int MINW, bool GEN>
__global__ void __launch_bounds__(WGW * WAVE, MINW) k_stage_wg(StageArgs A) {
    __shared__ Q4 lds_N[WGW * WAVE];
    __shared__ Q4 lds_F[WGW * WAVE];
    int lo, hi, first, stride;
    item_range(A.nitems, lo, hi, first, stride);
    for (int i = first; i < hi; i += stride) {
        stage_wg_item<CART, FLUX, GEN>(A, ld_const(A.items, i), lds_N, lds_F);
        __syncthreads();  // the next item reuses the exchange buffers
    }
}

// ---------------------------------------------------------------------------------------------
// paired stage: two waves per work item (the fused kernel's strip and rows, one halo row below).
// Wave H reconstructs variables 2H and 2H+1 (the variables of the reconstruction are independent);
// per row the waves exchange through LDS the halves of the face states the other needs (E and W
// go to wave 0, N and S to wave 1), wave 0 evaluates the x-face flux and wave 1 the y-face flux,
// the flux halves are exchanged, and each wave completes the residual and every Butcher output of
// its two variables.  On the last stage of a step wave 1 hands its half of the new state to
// wave 0, which evaluates the CFL quantity and realizability.  Same arithmetic, same association
// order as stage_item: each variable and each face flux is computed exactly as there.  Both waves
// run the same rows, so they reach the same barriers; every LDS slot is written after a barrier
// that follows the other wave's last read of it.
// ---------------------------------------------------------------------------------------------
constexpr int PAIR_WAVES = 2;

struct PairLds {
    double x1[PAIR_WAVES][4][WAVE];  // face-state halves: wave 0 writes N, S (vars 0-1); wave 1 E, W (vars 2-3)
    double x2[PAIR_WAVES][4][WAVE];  // flux halves: wave 0 writes FW, FE (vars 2-3); wave 1 F (vars 0-1)
    double x3[2][WAVE];              // wave 1's half (vars 2-3) of the new state, last stage
};

template <int H>
__device__ __forceinline__ double hv(const Q4& q, int i) { return q.v[2 * H + i]; }

// the two variables of this wave of every output of the stage for cell (kk, j); o = the state
// output's values of these variables
template <bool CART, int H>
__device__ __forceinline__ void emit_half(const StageArgs& A, const DevBlock& B, int kk, int j, double dt,
                                          const double res[2], double o[2]) {
    const long idx = B.off + 4 * ((long)kk * B.nx + j) + 2 * H;
    if (A.Rout) *reinterpret_cast<double2*>(A.Rout + idx) = make_double2(res[0], res[1]);
#pragma unroll 1
    for (int m0 = 0; m0 < A.nout; m0 += EMIT_CHUNK) {
        double2 s[EMIT_CHUNK];
#pragma unroll
        for (int q = 0; q < EMIT_CHUNK; ++q)
            if (m0 + q < A.nout) s[q] = *reinterpret_cast<const double2*>(A.src[m0 + q] + idx);
#pragma unroll
        for (int q = 0; q < EMIT_CHUNK; ++q) {
            const int m = m0 + q;
            if (m >= A.nout) break;
            const double am = A.a[m];
            double2 r = s[q];
            if (am != 0.0) {
                const double cdt = dt * am;  // dt * a[stage][step] (explicit_runge_kutta.py:72)
                r.x = s[q].x + cdt * res[0];
                r.y = s[q].y + cdt * res[1];
            }
            *reinterpret_cast<double2*>(A.dst[m] + idx) = r;
            if (m == A.state_out) { o[0] = r.x; o[1] = r.y; }
        }
    }
}

template <bool CART, int FLUX, bool GEN, int H>
__device__ __forceinline__ void stage_pair_body(const StageArgs& A, const WorkItem it, PairLds& L) {
    const DevBlock& B = const_ref(A.blocks, it.block);
    const int lane = threadIdx.x & (WAVE - 1);
    const int nx = B.nx, ny = B.ny;
    const int j = it.j0 - HALO + lane;
    const bool valid = (j >= -1) && (j <= nx);
    const bool inner = (j >= 0) && (j < nx);
    const bool gcol = (j == -1) || (j == nx);
    const bool recon = inner && (j >= it.j0 - 1) && (j <= it.j0 + STRIP);
    const bool xface = (j >= it.j0) && (j <= min(it.j0 + STRIP, nx));
    const bool owner = inner && (j >= it.j0) && (j < it.j0 + STRIP);
    const double dt = (A.dt != nullptr) ? ld_const(A.dt, B.case_id) : 0.0;
    const Phys ph = A.ph;
    const bool cfl = A.dtmin != nullptr && A.state_out >= 0;

    const int kr0 = max(it.r0 - 1, 0);
    const int kr1 = min(it.r1, ny);  // row ny is a virtual row closing the north boundary face

    Q4 Um = phx::q4zero(), Uc = phx::q4zero(), Up = phx::q4zero();
    if (inner) Um = fetch_r<CART, GEN>(A.blocks, B, A.Uin, kr0 - 1, j, ph);
    if (valid) Uc = fetch_r<CART, GEN>(A.blocks, B, A.Uin, kr0, j, ph);

    Q4 QNprev = phx::q4zero();  // wave 1: N state of the previous row
    double part[2] = {0.0, 0.0}, prevLN = 0.0, prevA = 1.0;
    double lane_dt = __longlong_as_double((long long)DT_INF_BITS);
    bool lane_bad = false;

    for (int k = kr0; k <= kr1; ++k) {
        const bool real = (k < ny);
        if (real) {
            if (inner) Up = fetch_r<CART, GEN>(A.blocks, B, A.Uin, k + 1, j, ph);
            else if (gcol && k + 1 <= kr1 && k + 1 < ny) Up = fetch_r<CART, GEN>(A.blocks, B, A.Uin, k + 1, j, ph);
        }
        // reconstruction of this wave's two variables
        double e[2] = {0.0, 0.0}, w[2] = {0.0, 0.0}, n[2] = {0.0, 0.0}, s[2] = {0.0, 0.0};
        {
            double uw[2], ue[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                uw[i] = __shfl_up(hv<H>(Uc, i), 1, WAVE);
                ue[i] = __shfl_down(hv<H>(Uc, i), 1, WAVE);
            }
            if (real && recon) {
                const phx::CellGeo g = cell_geo<CART>(B, k, j);
#pragma unroll
                for (int i = 0; i < 2; ++i)
                    phx::reconstruct_var<GEN>(hv<H>(Uc, i), uw[i], ue[i], hv<H>(Um, i), hv<H>(Up, i), g, ph.limiter,
                                              e[i], w[i], n[i], s[i]);
            }
        }
        // exchange 1: wave 0 needs E and W of variables 2-3, wave 1 N and S of variables 0-1
        if (H == 0) {
            L.x1[0][0][lane] = n[0]; L.x1[0][1][lane] = n[1];
            L.x1[0][2][lane] = s[0]; L.x1[0][3][lane] = s[1];
        } else {
            L.x1[1][0][lane] = e[0]; L.x1[1][1][lane] = e[1];
            L.x1[1][2][lane] = w[0]; L.x1[1][3][lane] = w[1];
        }
        __syncthreads();
        const bool xrow = (k >= it.r0) && (k < it.r1);
        double fw[2] = {0.0, 0.0}, fe[2] = {0.0, 0.0}, fy[2] = {0.0, 0.0};
        if (H == 0) {
            Q4 rE, rW;
            rE.v[0] = e[0]; rE.v[1] = e[1]; rE.v[2] = L.x1[1][0][lane]; rE.v[3] = L.x1[1][1][lane];
            rW.v[0] = w[0]; rW.v[1] = w[1]; rW.v[2] = L.x1[1][2][lane]; rW.v[3] = L.x1[1][3][lane];
            if (!(real && recon)) rE = rW = phx::q4zero();
            const Q4 qsend = (j == -1) ? Uc : rE;
            const Q4 QEl = shfl_up1(qsend);  // limited east state of cell j-1 (ghost value for j == 0)
            Q4 FW = phx::q4zero();
            if (xrow && xface) {
                const Q4 Lx = (j == 0) ? outer_state<CART>(B, PHX_W, rW, QEl, k, j) : QEl;
                const Q4 Rx = (j == nx) ? outer_state<CART>(B, PHX_E, QEl, Uc, k, j) : rW;
                FW = xflux<CART, FLUX, GEN>(B, Lx, Rx, k, j, ph);
            }
            const Q4 FE = shfl_down1(FW);
            // exchange 2 (write): variables 2-3 of FW and FE for wave 1
            L.x2[0][0][lane] = FW.v[2]; L.x2[0][1][lane] = FW.v[3];
            L.x2[0][2][lane] = FE.v[2]; L.x2[0][3][lane] = FE.v[3];
            fw[0] = FW.v[0]; fw[1] = FW.v[1];
            fe[0] = FE.v[0]; fe[1] = FE.v[1];
        } else {
            Q4 rN, rS;
            rN.v[0] = L.x1[0][0][lane]; rN.v[1] = L.x1[0][1][lane]; rN.v[2] = n[0]; rN.v[3] = n[1];
            rS.v[0] = L.x1[0][2][lane]; rS.v[1] = L.x1[0][3][lane]; rS.v[2] = s[0]; rS.v[3] = s[1];
            if (!(real && recon)) rN = rS = phx::q4zero();
            Q4 F = phx::q4zero();
            if (k >= it.r0 && owner) {
                // face between rows k-1 and k (south face of row k, north face of row k-1)
                const Q4 Ly = (k == 0) ? outer_state<CART>(B, PHX_S, rS, Um, k, j) : QNprev;
                const Q4 Ry = (k == ny) ? outer_state<CART>(B, PHX_N, QNprev, Uc, k - 1, j) : rS;
                F = yflux<CART, FLUX, GEN>(B, Ly, Ry, k, j, ph);
            }
            // exchange 2 (write): variables 0-1 of F for wave 0
            L.x2[1][0][lane] = F.v[0]; L.x2[1][1][lane] = F.v[1];
            fy[0] = F.v[2]; fy[1] = F.v[3];
            QNprev = rN;
        }
        __syncthreads();
        if (H == 0) {
            fy[0] = L.x2[1][0][lane]; fy[1] = L.x2[1][1][lane];
        } else {
            fw[0] = L.x2[0][0][lane]; fw[1] = L.x2[0][1][lane];
            fe[0] = L.x2[0][2][lane]; fe[1] = L.x2[0][3][lane];
        }
        double o[2] = {0.0, 0.0};
        const bool emit = (k >= it.r0) && owner && (k - 1 >= it.r0);
        if (k >= it.r0 && owner) {
            if (emit) {
                // residual of row k-1: 0.5*(W - E + S - N)/A with I = L*(2*F) (fvm/base.py:141-190)
                double res[2];
#pragma unroll
                for (int i = 0; i < 2; ++i) res[i] = (0.5 * (part[i] - prevLN * (2.0 * fy[i]))) / prevA;
                emit_half<CART, H>(A, B, k - 1, j, dt, res, o);
            }
            if (xrow) {
                double LW, LE, LS, LN, Ac;
                cell_L<CART>(B, k, j, LW, LE, LS, LN, Ac);
#pragma unroll
                for (int i = 0; i < 2; ++i) part[i] = (LW * (2.0 * fw[i]) - LE * (2.0 * fe[i])) + LS * (2.0 * fy[i]);
                prevLN = LN;
                prevA = Ac;
            }
        }
        if (cfl) {  // uniform: the new state's CFL quantity and realizability on wave 0
            if (H == 1) { L.x3[0][lane] = o[0]; L.x3[1][lane] = o[1]; }
            __syncthreads();
            if (H == 0 && emit) {
                Q4 q;
                q.v[0] = o[0]; q.v[1] = o[1]; q.v[2] = L.x3[0][lane]; q.v[3] = L.x3[1][lane];
                double dx, dy;
                cell_dxdy<CART>(B, k - 1, j, dx, dy);
                const double t = phx::cell_dt(q, dx, dy, ph);
                if (!(q.v[0] > 0.0) || !(q.v[3] > 0.0) || !(t == t)) lane_bad = true;
                else lane_dt = fmin(lane_dt, t);
            }
        }
        Um = Uc;
        Uc = Up;
    }
    if (H == 0) reduce_item(A, B, lane, lane_dt, lane_bad);
}

template <bool CART, int FLUX, int MINW, bool GEN>
__global__ void __launch_bounds__(PAIR_WAVES * WAVE, MINW) k_stage_pair(StageArgs A) {
    __shared__ PairLds L;
    const int h = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    int lo, hi, first, stride;
    item_range(A.nitems, lo, hi, first, stride);
    for (int i = first; i < hi; i += stride) {
        const WorkItem it = ld_const(A.items, i);
        const DevBlock& B = const_ref(A.blocks, it.block);
        if (A.skip_dead && A.bad[B.case_id] < *A.step) continue;  // uniform over the workgroup
        if (h == 0) stage_pair_body<CART, FLUX, GEN, 0>(A, it, L);
        else stage_pair_body<CART, FLUX, GEN, 1>(A, it, L);
        __syncthreads();  // the next item reuses the exchange buffers
    }
}

// ---------------------------------------------------------------------------------------------
// split stage: for small meshes, where short rolling-row items would spend most of their work
// on the halo rows they reconstruct again, the stage runs as three kernels: every cell is
// reconstructed exactly once (k_recon_x, with the x-faces), every y-face flux is evaluated once
// from the stored N/S states (k_yface, one lane per face), then the residual and Butcher updates
// (k_update).  Same arithmetic, same association order as stage_item.
// ---------------------------------------------------------------------------------------------
// k_recon_x: owned columns per wave.  Two halo lanes per side as in the fused kernel: the halo
// cell j0-1 is reconstructed (its E state feeds the W face of the first owned cell) and needs its
// own west neighbour j0-2; likewise on the east side.
constexpr int STRIP_A = WAVE - 2 * HALO;
constexpr int STRIP_B = WAVE;      // k_yface, k_update: no column halo

// ---- nq-point Gauss faces (split stage only; GEN instantiation, nq = A.ph.nq in 2..QMAX) ----
// Face F's sum S_F = sum_i w_i*F_i starting from 0 (integrate_flux, fvm/base.py:168-190); per cell
// the split buffers hold p1 = LW*S_W - LE*S_E, the nq N states then the nq S states
// (8*nq doubles at 2*nq*off + 8*nq*cell) and per y-face S_F.
__device__ __forceinline__ void acc_w(Q4& acc, double w, const Q4& F) {
#pragma unroll
    for (int c = 0; c < 4; ++c) acc.v[c] = acc.v[c] + w * F.v[c];
}

template <bool CART, int FLUX>
__device__ __forceinline__ void recon_x_item_q(const StageArgs& A, const WorkItem it) {
    const DevBlock& B = A.blocks[it.block];
    if (A.skip_dead && A.bad[B.case_id] < *A.step) return;
    const int nx = B.nx;
    const int j = it.j0 - HALO + threadIdx.x % WAVE;
    const bool valid = (j >= -1) && (j <= nx);
    const bool inner = (j >= 0) && (j < nx);
    const bool gcol = (j == -1) || (j == nx);
    const bool recon = inner && (j >= it.j0 - 1) && (j <= it.j0 + STRIP_A);
    const bool xface = (j >= it.j0) && (j <= min(it.j0 + STRIP_A, nx));
    const bool owner = inner && (j >= it.j0) && (j < it.j0 + STRIP_A);
    const Phys ph = A.ph;
    const int nq = ph.nq;

    Q4 Um = phx::q4zero(), Uc = phx::q4zero(), Up = phx::q4zero();
    if (inner) Um = fetch_r<CART, true>(A.blocks, B, A.Uin, it.r0 - 1, j, ph);
    if (valid) Uc = fetch_r<CART, true>(A.blocks, B, A.Uin, it.r0, j, ph);
    for (int k = it.r0; k < it.r1; ++k) {
        if (inner) Up = fetch_r<CART, true>(A.blocks, B, A.Uin, k + 1, j, ph);
        else if (gcol && k + 1 < it.r1) Up = fetch_r<CART, true>(A.blocks, B, A.Uin, k + 1, j, ph);
        const Q4 Uw = shfl_up1(Uc);
        const Q4 Ue = shfl_down1(Uc);
        phx::ReconQ r;
        r.u = r.gx = r.gy = r.phi = phx::q4zero();
        phx::QGeo qg;
        for (int f = 0; f < 4; ++f)
            for (int i = 0; i < phx::QMAX; ++i) qg.dx[f][i] = qg.dy[f][i] = 0.0;
        if (recon) {
            const phx::CellGeo g = cell_geo<CART>(B, k, j);
            qg = cell_qgeo(B, k, j, nq);
            r = phx::reconstruct_q(Uc, Uw, Ue, Um, Up, g, qg, nq, ph.limiter);
        }
        Q4 SW = phx::q4zero();  // sum over the W face's points of w_i * F_i
#pragma unroll
        for (int i = 0; i < phx::QMAX; ++i) {
            if (i >= nq) break;
            const Q4 qsend = (j == -1) ? Uc : phx::qstate(r, qg, PHX_E, i);
            const Q4 QEl = shfl_up1(qsend);  // limited east state of cell j-1 at point i
            if (xface) {
                const Q4 Wi = phx::qstate(r, qg, PHX_W, i);
                const Q4 L = (j == 0) ? outer_state<CART>(B, PHX_W, Wi, QEl, k, j) : QEl;
                const Q4 R = (j == nx) ? outer_state<CART>(B, PHX_E, QEl, Uc, k, j) : Wi;
                acc_w(SW, phx::gauss_w(nq, i), xflux<CART, FLUX, true>(B, L, R, k, j, ph));
            }
        }
        const Q4 SE = shfl_down1(SW);
        if (owner) {
            double LW, LE, LS, LN, Ac;
            cell_L<CART>(B, k, j, LW, LE, LS, LN, Ac);
            Q4 part;
#pragma unroll
            for (int c = 0; c < 4; ++c) part.v[c] = LW * SW.v[c] - LE * SE.v[c];
            const long cell = (long)k * nx + j;
            st4(A.p1 + B.off + 4 * cell, part);
            double* sn = A.sn + 2 * nq * B.off + 8 * nq * cell;
#pragma unroll
            for (int i = 0; i < phx::QMAX; ++i) {
                if (i >= nq) break;
                st4(sn + 4 * i, phx::qstate(r, qg, PHX_N, i));
                st4(sn + 4 * (nq + i), phx::qstate(r, qg, PHX_S, i));
            }
        }
        Um = Uc;
        Uc = Up;
    }
}

// Gauss sum S_F = sum_i w_i F_i (from 0) of y-face f of column j
template <bool CART, int FLUX>
__device__ __forceinline__ Q4 yface_sum_q(const StageArgs& A, const DevBlock& B, const double* sn, int f, int j, const Phys& ph) {
    const int nx = B.nx, ny = B.ny;
    const int nq = ph.nq;
    // N states of row f-1 (f > 0) and S states of row f (f < ny)
    const long below = f > 0 ? 8 * nq * ((long)(f - 1) * nx + j) : 0;
    const long above = f < ny ? 8 * nq * ((long)f * nx + j) + 4 * nq : 0;
    Q4 gS = phx::q4zero(), gN = phx::q4zero();
    if (f == 0) gS = fetch_r<CART, true>(A.blocks, B, A.Uin, -1, j, ph);
    if (f == ny) gN = fetch_r<CART, true>(A.blocks, B, A.Uin, ny, j, ph);
    Q4 S = phx::q4zero();
#pragma unroll
    for (int i = 0; i < phx::QMAX; ++i) {
        if (i >= nq) break;
        Q4 L, R;
        if (f == 0) L = outer_state<CART>(B, PHX_S, ld4(sn + above + 4 * i), gS, 0, j);
        else L = ld4(sn + below + 4 * i);
        if (f == ny) R = outer_state<CART>(B, PHX_N, L, gN, ny - 1, j);
        else R = ld4(sn + above + 4 * i);
        acc_w(S, phx::gauss_w(nq, i), yflux<CART, FLUX, true>(B, L, R, f, j, ph));
    }
    return S;
}

template <bool CART, int FLUX>
__device__ __forceinline__ void yface_item_q(const StageArgs& A, const WorkItem it) {
    const DevBlock& B = A.blocks[it.block];
    if (A.skip_dead && A.bad[B.case_id] < *A.step) return;
    const int nx = B.nx;
    const int j = it.j0 + (int)(threadIdx.x % WAVE);
    if (j >= nx) return;
    const double* sn = A.sn + 2 * A.ph.nq * B.off;
    for (int f = it.r0; f < it.r1; ++f)
        st4(A.yf + B.yoff + 4 * ((long)f * nx + j), yface_sum_q<CART, FLUX>(A, B, sn, f, j, A.ph));
}

// Gauss faces, y-faces and update in one pass (variant 5; yface_update_item with the sums S_F)
template <bool CART, int FLUX>
__device__ __forceinline__ void yface_update_item_q(const StageArgs& A, const WorkItem it) {
    const DevBlock& B = A.blocks[it.block];
    if (A.skip_dead && A.bad[B.case_id] < *A.step) return;
    const int lane = threadIdx.x & (WAVE - 1);
    const int nx = B.nx;
    const int j = it.j0 + lane;
    const double dt = (A.dt != nullptr) ? ld_const(A.dt, B.case_id) : 0.0;
    const Phys ph = A.ph;
    double lane_dt = __longlong_as_double((long long)DT_INF_BITS);
    bool lane_bad = false;
    if (j < nx) {
        const double* sn = A.sn + 2 * ph.nq * B.off;
        Q4 FS = yface_sum_q<CART, FLUX>(A, B, sn, it.r0, j, ph);
        for (int k = it.r0; k < it.r1; ++k) {
            const Q4 FN = yface_sum_q<CART, FLUX>(A, B, sn, k + 1, j, ph);
            const Q4 p = ld4(A.p1 + B.off + 4 * ((long)k * nx + j));
            double LW, LE, LS, LN, Ac;
            cell_L<CART>(B, k, j, LW, LE, LS, LN, Ac);
            Q4 Res;
#pragma unroll
            for (int c = 0; c < 4; ++c) Res.v[c] = (0.5 * ((p.v[c] + LS * FS.v[c]) - LN * FN.v[c])) / Ac;
            emit_row<CART>(A, B, k, j, dt, Res, lane_dt, lane_bad);
            FS = FN;
        }
    }
    reduce_item(A, B, lane, lane_dt, lane_bad);
}

template <bool CART, int FLUX, bool GEN>
__device__ __forceinline__ void recon_x_item(const StageArgs& A, const WorkItem it) {
    const DevBlock& B = const_ref(A.blocks, it.block);
    if (A.skip_dead && A.bad[B.case_id] < *A.step) return;
    const int nx = B.nx;
    const int j = it.j0 - HALO + threadIdx.x % WAVE;
    const bool valid = (j >= -1) && (j <= nx);
    const bool inner = (j >= 0) && (j < nx);
    const bool gcol = (j == -1) || (j == nx);
    const bool recon = inner && (j >= it.j0 - 1) && (j <= it.j0 + STRIP_A);
    const bool xface = (j >= it.j0) && (j <= min(it.j0 + STRIP_A, nx));
    const bool owner = inner && (j >= it.j0) && (j < it.j0 + STRIP_A);
    const Phys ph = A.ph;

    Q4 Um = phx::q4zero(), Uc = phx::q4zero(), Up = phx::q4zero();
    if (inner) Um = fetch_r<CART, GEN>(A.blocks, B, A.Uin, it.r0 - 1, j, ph);
    if (valid) Uc = fetch_r<CART, GEN>(A.blocks, B, A.Uin, it.r0, j, ph);
    for (int k = it.r0; k < it.r1; ++k) {
        if (inner) Up = fetch_r<CART, GEN>(A.blocks, B, A.Uin, k + 1, j, ph);
        else if (gcol && k + 1 < it.r1) Up = fetch_r<CART, GEN>(A.blocks, B, A.Uin, k + 1, j, ph);
        const Q4 Uw = shfl_up1(Uc);
        const Q4 Ue = shfl_down1(Uc);
        phx::Recon r;
        r.E = r.W = r.N = r.S = phx::q4zero();
        if (recon) {
            const phx::CellGeo g = cell_geo<CART>(B, k, j);
            r = phx::reconstruct<GEN>(Uc, Uw, Ue, Um, Up, g, ph.limiter);
        }
        const Q4 qsend = (j == -1) ? Uc : r.E;
        const Q4 QEl = shfl_up1(qsend);  // limited east state of cell j-1 (ghost value for j == 0)
        Q4 FW = phx::q4zero();
        if (xface) {
            const Q4 L = (j == 0) ? outer_state<CART>(B, PHX_W, r.W, QEl, k, j) : QEl;
            const Q4 R = (j == nx) ? outer_state<CART>(B, PHX_E, QEl, Uc, k, j) : r.W;
            FW = xflux<CART, FLUX, GEN>(B, L, R, k, j, ph);
        }
        const Q4 FE = shfl_down1(FW);
        if (owner) {
            double LW, LE, LS, LN, Ac;
            cell_L<CART>(B, k, j, LW, LE, LS, LN, Ac);
            Q4 part;
#pragma unroll
            for (int c = 0; c < 4; ++c) part.v[c] = LW * (2.0 * FW.v[c]) - LE * (2.0 * FE.v[c]);
            const long cell = (long)k * nx + j;
            st4(A.p1 + B.off + 4 * cell, part);
            st4(A.sn + 2 * B.off + 8 * cell, r.N);
            st4(A.sn + 2 * B.off + 8 * cell + 4, r.S);
        }
        Um = Uc;
        Uc = Up;
    }
}

// y-faces of face rows [r0, r1) of a strip of 64 columns (face f between rows f-1 and f; f = 0 and
// f = ny take the outer states of the S and N boundaries): one flux per lane and face, stored.
template <bool CART, int FLUX, bool GEN>
__device__ __forceinline__ void yface_item(const StageArgs& A, const WorkItem it) {
    const DevBlock& B = A.blocks[it.block];
    if (A.skip_dead && A.bad[B.case_id] < *A.step) return;
    const int nx = B.nx, ny = B.ny;
    const int j = it.j0 + (int)(threadIdx.x % WAVE);
    if (j >= nx) return;
    const Phys ph = A.ph;
    const double* sn = A.sn + 2 * B.off;
    for (int f = it.r0; f < it.r1; ++f) {
        Q4 L, R;
        if (f == 0) L = outer_state<CART>(B, PHX_S, ld4(sn + 8 * (long)j + 4), fetch_r<CART, GEN>(A.blocks, B, A.Uin, -1, j, ph), 0, j);
        else L = ld4(sn + 8 * ((long)(f - 1) * nx + j));
        if (f == ny) R = outer_state<CART>(B, PHX_N, L, fetch_r<CART, GEN>(A.blocks, B, A.Uin, ny, j, ph), ny - 1, j);
        else R = ld4(sn + 8 * ((long)f * nx + j) + 4);
        st4(A.yf + B.yoff + 4 * ((long)f * nx + j), yflux<CART, FLUX, GEN>(B, L, R, f, j, ph));
    }
}

// residual ((x-part + LS*(2 F_S)) - LN*(2 F_N))/A (Gauss sums S_F in place of 2 F with nq > 1) of rows [r0, r1) of a 64-column strip, then every
// Butcher output, the CFL quantity and realizability (as the fused kernel).
template <bool CART>
__device__ __forceinline__ void update_item(const StageArgs& A, const WorkItem it) {
    const DevBlock& B = A.blocks[it.block];
    if (A.skip_dead && A.bad[B.case_id] < *A.step) return;
    const int lane = threadIdx.x & (WAVE - 1);
    const int nx = B.nx;
    const int j = it.j0 + lane;
    const double dt = (A.dt != nullptr) ? ld_const(A.dt, B.case_id) : 0.0;
    double lane_dt = __longlong_as_double((long long)DT_INF_BITS);
    bool lane_bad = false;
    if (j < nx) {
        const double* yf = A.yf + B.yoff + 4 * (long)j;
        Q4 FS = ld4(yf + 4 * (long)it.r0 * nx);
        for (int k = it.r0; k < it.r1; ++k) {
            const Q4 FN = ld4(yf + 4 * (long)(k + 1) * nx);
            const Q4 p = ld4(A.p1 + B.off + 4 * ((long)k * nx + j));
            double LW, LE, LS, LN, Ac;
            cell_L<CART>(B, k, j, LW, LE, LS, LN, Ac);
            Q4 Res;
            if (A.ph.nq > 1) {  // y-face entries are the Gauss sums already
#pragma unroll
                for (int c = 0; c < 4; ++c) Res.v[c] = (0.5 * ((p.v[c] + LS * FS.v[c]) - LN * FN.v[c])) / Ac;
            } else {
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    Res.v[c] = (0.5 * ((p.v[c] + LS * (2.0 * FS.v[c])) - LN * (2.0 * FN.v[c]))) / Ac;
            }
            emit_row<CART>(A, B, k, j, dt, Res, lane_dt, lane_bad);
            FS = FN;
        }
    }
    reduce_item(A, B, lane, lane_dt, lane_bad);
}

// y-faces and update in one pass (variant 5): the lane of column j walks the face rows of cell
// rows [r0, r1); after the flux of face k+1 it has both y-fluxes of cell (k, j) and completes its
// residual and Butcher outputs as update_item does.  Face r0 is evaluated again by the item
// below (one extra flux per item and column instead of a y-flux buffer and a third kernel).
template <bool CART, int FLUX, bool GEN>
__device__ __forceinline__ Q4 yface_flux_at(const StageArgs& A, const DevBlock& B, const double* sn, int f, int j,
                                            const Phys& ph) {
    const int nx = B.nx, ny = B.ny;
    Q4 L, R;
    if (f == 0) L = outer_state<CART>(B, PHX_S, ld4(sn + 8 * (long)j + 4), fetch_r<CART, GEN>(A.blocks, B, A.Uin, -1, j, ph), 0, j);
    else L = ld4(sn + 8 * ((long)(f - 1) * nx + j));
    if (f == ny) R = outer_state<CART>(B, PHX_N, L, fetch_r<CART, GEN>(A.blocks, B, A.Uin, ny, j, ph), ny - 1, j);
    else R = ld4(sn + 8 * ((long)f * nx + j) + 4);
    return yflux<CART, FLUX, GEN>(B, L, R, f, j, ph);
}

template <bool CART, int FLUX, bool GEN>
__device__ __forceinline__ void yface_update_item(const StageArgs& A, const WorkItem it) {
    const DevBlock& B = A.blocks[it.block];
    if (A.skip_dead && A.bad[B.case_id] < *A.step) return;
    const int lane = threadIdx.x & (WAVE - 1);
    const int nx = B.nx;
    const int j = it.j0 + lane;
    const double dt = (A.dt != nullptr) ? ld_const(A.dt, B.case_id) : 0.0;
    const Phys ph = A.ph;
    double lane_dt = __longlong_as_double((long long)DT_INF_BITS);
    bool lane_bad = false;
    if (j < nx) {
        const double* sn = A.sn + 2 * B.off;
        Q4 FS = yface_flux_at<CART, FLUX, GEN>(A, B, sn, it.r0, j, ph);
        for (int k = it.r0; k < it.r1; ++k) {
            const Q4 FN = yface_flux_at<CART, FLUX, GEN>(A, B, sn, k + 1, j, ph);
            const Q4 p = ld4(A.p1 + B.off + 4 * ((long)k * nx + j));
            double LW, LE, LS, LN, Ac;
            cell_L<CART>(B, k, j, LW, LE, LS, LN, Ac);
            Q4 Res;
#pragma unroll
            for (int c = 0; c < 4; ++c)
                Res.v[c] = (0.5 * ((p.v[c] + LS * (2.0 * FS.v[c])) - LN * (2.0 * FN.v[c]))) / Ac;
            emit_row<CART>(A, B, k, j, dt, Res, lane_dt, lane_bad);
            FS = FN;
        }
    }
    reduce_item(A, B, lane, lane_dt, lane_bad);
}

// y-faces and update with one face per wave (variant 6): a workgroup of R+1 waves takes R cell rows
// [r0, r1) of a 64-column strip; wave w evaluates face r0+w into LDS, and after one barrier wave w < R
// completes cell row r0+w from faces r0+w and r0+w+1 as update_item does.  Every wave's chain is one
// flux and at most one update, and a face is evaluated twice only at item boundaries.
constexpr int YU_MAX_ROWS = 7;  // rows per item (R + 1 <= 8 waves per workgroup)

// Q = true: Gauss faces (the sums S_F of yface_sum_q, the residual of the nq-point rule)
template <bool CART, int FLUX, bool GEN, bool Q>
__device__ __forceinline__ void yface_update_wg_item(const StageArgs& A, const WorkItem it, Q4* lds) {
    const DevBlock B = ld_const(A.blocks, it.block);  // descriptor in scalar registers
    if (A.skip_dead && A.bad[B.case_id] < *A.step) return;  // uniform over the workgroup
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = threadIdx.x / WAVE;
    const int nx = B.nx;
    const int j = it.j0 + lane;
    const int nrow = it.r1 - it.r0;
    const Phys ph = A.ph;
    if (w <= nrow && j < nx) {
        if (Q) lds[w * WAVE + lane] = yface_sum_q<CART, FLUX>(A, B, A.sn + 2 * ph.nq * B.off, it.r0 + w, j, ph);
        else lds[w * WAVE + lane] = yface_flux_at<CART, FLUX, GEN>(A, B, A.sn + 2 * B.off, it.r0 + w, j, ph);
    }
    __syncthreads();
    double lane_dt = __longlong_as_double((long long)DT_INF_BITS);
    bool lane_bad = false;
    if (w < nrow && j < nx) {
        const int k = it.r0 + w;
        const double dt = (A.dt != nullptr) ? ld_const(A.dt, B.case_id) : 0.0;
        const Q4 FS = lds[w * WAVE + lane], FN = lds[(w + 1) * WAVE + lane];
        const Q4 p = ld4(A.p1 + B.off + 4 * ((long)k * nx + j));
        double LW, LE, LS, LN, Ac;
        cell_L<CART>(B, k, j, LW, LE, LS, LN, Ac);
        Q4 Res;
        if (Q) {  // the face entries are the Gauss sums already
#pragma unroll
            for (int c = 0; c < 4; ++c) Res.v[c] = (0.5 * ((p.v[c] + LS * FS.v[c]) - LN * FN.v[c])) / Ac;
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c)
                Res.v[c] = (0.5 * ((p.v[c] + LS * (2.0 * FS.v[c])) - LN * (2.0 * FN.v[c]))) / Ac;
        }
        emit_row<CART>(A, B, k, j, dt, Res, lane_dt, lane_bad);
    }
    reduce_item(A, B, lane, lane_dt, lane_bad);
}

// CK: geometry kind of the launch's blocks, 0 general, 1 Cartesian, 2 both (the path chosen per
// item's block; one launch over the merged list of a mesh with both kinds)
template <int CK, int FLUX, bool GEN, bool Q>
__global__ void __launch_bounds__((YU_MAX_ROWS + 1) * WAVE) k_yface_update_wg(StageArgs A) {
    __shared__ Q4 lds[(YU_MAX_ROWS + 1) * WAVE];
    int lo, hi, first, stride;
    item_range(A.nitems, lo, hi, first, stride);
    for (int i = first; i < hi; i += stride) {
        const WorkItem it = ld_const(A.items, i);
        if (CK == 2) {
            if (ld_const(A.blocks, it.block).kind == PHX_GEOM_CARTESIAN) yface_update_wg_item<true, FLUX, GEN, Q>(A, it, lds);
            else yface_update_wg_item<false, FLUX, GEN, Q>(A, it, lds);
        } else {
            yface_update_wg_item<CK == 1, FLUX, GEN, Q>(A, it, lds);
        }
        __syncthreads();  // the next item reuses the face buffer
    }
}

// y-faces and update over flat cell ranges (variant 8): a workgroup of T lanes takes the cells
// c in [c0, c0 + n) of one block in row-major order.  The S face of cell c has the face index c and
// its N face c + nx (face (f, j) at f*nx + j, f = 0..ny), so the faces of the range are the n + nx
// consecutive indices from c0: lane t evaluates faces c0 + t, c0 + t + T, ... into LDS and, after
// one barrier, completes cells c0 + t, c0 + t + T, ... as update_item does.  No lane idles on a
// partial column strip, and a face is evaluated twice only where two ranges meet (nx faces per
// range).  T = (A.rows_wave + 1) * WAVE; the host sizes the ranges so that n + nx <= A.fq * T, the
// faces the dynamic LDS holds.
template <bool CART, int FLUX, bool GEN>
__device__ __forceinline__ void yface_update_flat_item(const StageArgs& A, const WorkItem it, Q4* lds) {
    const DevBlock B = ld_const(A.blocks, it.block);  // descriptor in scalar registers
    if (A.skip_dead && A.bad[B.case_id] < *A.step) return;  // uniform over the workgroup
    const int T = blockDim.x;
    const int t = threadIdx.x;
    const int lane = t & (WAVE - 1);
    const int nx = B.nx;
    const int c0 = it.j0, n = it.r0;
    const Phys ph = A.ph;
    const double* sn = A.sn + 2 * B.off;
    for (int q = t; q < n + nx; q += T) {
        const int fi = c0 + q;
        const int f = fi / nx;
        lds[q] = yface_flux_at<CART, FLUX, GEN>(A, B, sn, f, fi - f * nx, ph);
    }
    __syncthreads();
    double lane_dt = __longlong_as_double((long long)DT_INF_BITS);
    bool lane_bad = false;
    const double dt = (A.dt != nullptr) ? ld_const(A.dt, B.case_id) : 0.0;
    for (int q = t; q < n; q += T) {
        const int c = c0 + q;
        const int k = c / nx, j = c - k * nx;
        const Q4 FS = lds[q], FN = lds[q + nx];
        const Q4 p = ld4(A.p1 + B.off + 4 * (long)c);
        double LW, LE, LS, LN, Ac;
        cell_L<CART>(B, k, j, LW, LE, LS, LN, Ac);
        Q4 Res;
#pragma unroll
        for (int cc = 0; cc < 4; ++cc)
            Res.v[cc] = (0.5 * ((p.v[cc] + LS * (2.0 * FS.v[cc])) - LN * (2.0 * FN.v[cc]))) / Ac;
        emit_row<CART>(A, B, k, j, dt, Res, lane_dt, lane_bad);
    }
    reduce_item(A, B, lane, lane_dt, lane_bad);
}

constexpr int FLAT_MAX_WAVES = 16;  // workgroup of up to 1024 lanes

template <bool CART, int FLUX, bool GEN>
__global__ void __launch_bounds__(FLAT_MAX_WAVES * WAVE) k_yface_update_flat(StageArgs A) {
    extern __shared__ Q4 flat_lds[];
    int lo, hi, first, stride;
    item_range(A.nitems, lo, hi, first, stride);
    for (int i = first; i < hi; i += stride) {
        yface_update_flat_item<CART, FLUX, GEN>(A, ld_const(A.items, i), flat_lds);
        __syncthreads();  // the next item reuses the face buffer
    }
}

template <bool CART, int FLUX, bool GEN>
__global__ void __launch_bounds__(WAVE) k_yface_update(StageArgs A) {
    int lo, hi, first, stride;
    item_range(A.nitems, lo, hi, first, stride);
    for (int i = first; i < hi; i += stride) yface_update_item<CART, FLUX, GEN>(A, ld_const(A.items, i));
}

template <int CK, int FLUX, int MINW, bool GEN>
__global__ void __launch_bounds__(WAVE, MINW) k_recon_x(StageArgs A) {
    int lo, hi, first, stride;
    item_range(A.nitems, lo, hi, first, stride);
    for (int i = first; i < hi; i += stride) {
        const WorkItem it = ld_const(A.items, i);
        if (CK == 2) {
            if (ld_const(A.blocks, it.block).kind == PHX_GEOM_CARTESIAN) recon_x_item<true, FLUX, GEN>(A, it);
            else recon_x_item<false, FLUX, GEN>(A, it);
        } else {
            recon_x_item<CK == 1, FLUX, GEN>(A, it);
        }
    }
}

// k_recon_x_lin: the reconstruction kernel of the split stage over the row-major positions of a
// block, p = k*(nx+1) + j with j in [0, nx]: j < nx is cell (k, j) and its west face, j = nx the
// east boundary face of row k.  Lane l of an item takes position j0 - 1 + l; lane 0 only
// reconstructs (its limited east state feeds the west face of lane 1), lane WAVE-1 only evaluates
// its face (the east face flux of lane WAVE-2), lanes 1..WAVE-2 own their positions.  East and west
// neighbour values come from the adjacent lanes inside a row and from memory at row ends and at the
// wave's two edge lanes.  Rows that are not a multiple of a column strip leave no idle lanes, and
// each wave needs one halo lane per side instead of two.  Same arithmetic and association order as
// recon_x_item.
constexpr int LIN_OWN = WAVE - 2;

template <bool CART, int FLUX, bool GEN>
__device__ __forceinline__ void recon_x_lin_item(const StageArgs& A, const WorkItem it) {
    const DevBlock B = ld_const(A.blocks, it.block);  // descriptor in scalar registers
    if (A.skip_dead && A.bad[B.case_id] < *A.step) return;
    const int nx = B.nx, npos = B.ny * (nx + 1);
    const int lane = threadIdx.x % WAVE;
    const int p = it.j0 - 1 + lane;
    const bool valid = (p >= 0) && (p < npos);
    const int k = valid ? p / (nx + 1) : 0;
    const int j = valid ? p - k * (nx + 1) : 0;
    const bool cell = valid && (j < nx);
    const Phys ph = A.ph;

    // every load is issued before the first shuffle, so that a wave waits for memory once
    const bool wmem = cell && (j == 0 || lane == 0);
    const bool emem = cell && (lane == WAVE - 1);
    Q4 Uc = phx::q4zero(), Um = phx::q4zero(), Up = phx::q4zero(), Uw = phx::q4zero(), Ue = phx::q4zero();
    if (valid) Uc = fetch_r<CART, GEN>(A.blocks, B, A.Uin, k, j, ph);  // j == nx: the east ghost value
    if (cell) {
        Um = fetch_r<CART, GEN>(A.blocks, B, A.Uin, k - 1, j, ph);
        Up = fetch_r<CART, GEN>(A.blocks, B, A.Uin, k + 1, j, ph);
    }
    if (wmem) Uw = fetch_r<CART, GEN>(A.blocks, B, A.Uin, k, j - 1, ph);
    if (emem) Ue = fetch_r<CART, GEN>(A.blocks, B, A.Uin, k, j + 1, ph);
    const Q4 Uwl = shfl_up1(Uc);
    const Q4 Uel = shfl_down1(Uc);
    if (!wmem) Uw = Uwl;
    if (!emem) Ue = Uel;
    phx::Recon r;
    r.E = r.W = r.N = r.S = phx::q4zero();
    if (cell) {
        const phx::CellGeo g = cell_geo<CART>(B, k, j);
        r = phx::reconstruct<GEN>(Uc, Uw, Ue, Um, Up, g, ph.limiter);
    }
    const Q4 QEl = shfl_up1(r.E);  // limited east state of position p-1 (cell j-1 of the same row for j > 0)
    Q4 FW = phx::q4zero();
    if (valid && lane > 0) {
        const Q4 L = (j == 0) ? outer_state<CART>(B, PHX_W, r.W, Uw, k, j) : QEl;
        const Q4 R = (j == nx) ? outer_state<CART>(B, PHX_E, QEl, Uc, k, j) : r.W;
        FW = xflux<CART, FLUX, GEN>(B, L, R, k, j, ph);
    }
    const Q4 FE = shfl_down1(FW);
    if (cell && lane > 0 && lane < WAVE - 1) {
        double LW, LE, LS, LN, Ac;
        cell_L<CART>(B, k, j, LW, LE, LS, LN, Ac);
        Q4 part;
#pragma unroll
        for (int c = 0; c < 4; ++c) part.v[c] = LW * (2.0 * FW.v[c]) - LE * (2.0 * FE.v[c]);
        const long cidx = (long)k * nx + j;
        st4(A.p1 + B.off + 4 * cidx, part);
        st4(A.sn + 2 * B.off + 8 * cidx, r.N);
        st4(A.sn + 2 * B.off + 8 * cidx + 4, r.S);
    }
}

template <int CK, int FLUX, int MINW, bool GEN>
__global__ void __launch_bounds__(WAVE, MINW) k_recon_x_lin(StageArgs A) {
    int lo, hi, first, stride;
    item_range(A.nitems, lo, hi, first, stride);
    for (int i = first; i < hi; i += stride) {
        const WorkItem it = ld_const(A.items, i);
        if (CK == 2) {
            if (ld_const(A.blocks, it.block).kind == PHX_GEOM_CARTESIAN) recon_x_lin_item<true, FLUX, GEN>(A, it);
            else recon_x_lin_item<false, FLUX, GEN>(A, it);
        } else {
            recon_x_lin_item<CK == 1, FLUX, GEN>(A, it);
        }
    }
}

template <bool CART, int FLUX, bool GEN>
__global__ void __launch_bounds__(WAVE) k_yface(StageArgs A) {
    int lo, hi, first, stride;
    item_range(A.nitems, lo, hi, first, stride);
    for (int i = first; i < hi; i += stride) yface_item<CART, FLUX, GEN>(A, ld_const(A.items, i));
}

// Gauss faces (nq > 1): kernels of their own, so that the register budget of the 1-point kernels is
// not set by the nq-point path
template <bool CART, int FLUX>
__global__ void __launch_bounds__(WAVE, 2) k_recon_x_q(StageArgs A) {
    int lo, hi, first, stride;
    item_range(A.nitems, lo, hi, first, stride);
    for (int i = first; i < hi; i += stride) recon_x_item_q<CART, FLUX>(A, ld_const(A.items, i));
}

template <bool CART, int FLUX>
__global__ void __launch_bounds__(WAVE) k_yface_q(StageArgs A) {
    int lo, hi, first, stride;
    item_range(A.nitems, lo, hi, first, stride);
    for (int i = first; i < hi; i += stride) yface_item_q<CART, FLUX>(A, ld_const(A.items, i));
}

template <bool CART, int FLUX>
__global__ void __launch_bounds__(WAVE) k_yface_update_q(StageArgs A) {
    int lo, hi, first, stride;
    item_range(A.nitems, lo, hi, first, stride);
    for (int i = first; i < hi; i += stride) yface_update_item_q<CART, FLUX>(A, ld_const(A.items, i));
}

template <bool CART>
__global__ void __launch_bounds__(WAVE) k_update(StageArgs A) {
    int lo, hi, first, stride;
    item_range(A.nitems, lo, hi, first, stride);
    for (int i = first; i < hi; i += stride) update_item<CART>(A, ld_const(A.items, i));
}

// ---------------------------------------------------------------------------------------------
// element-wise stage (small meshes, where one lane per cell leaves most SIMDs with one or two
// waves): the same arithmetic as stage_item in three passes with more lanes in flight,
//   k_ew_recon   one lane per cell and variable: limited face states E, W, N, S (the variables of
//                the reconstruction are independent: reconstruct_var)
//   k_ew_face    one lane per face: left/right states (block-edge faces: outer state) -> flux
//   k_ew_update  one lane per cell: residual, every Butcher output, CFL quantity, realizability
// Tiles {block, first element, count}: EW_WG / 4 cells, EW_WG faces, EW_WG cells.  Cells are numbered
// k*nx + j in their block; faces as in the adjoint (x-faces (k, jf) first, then y-faces (kf, j)).
// ---------------------------------------------------------------------------------------------
constexpr int EW_WG = 256;
constexpr int EW_RCELLS = EW_WG / 4;

// variable v of cell (k, j) (one cell outside the block at most, never a corner), in the
// reconstruction variables; own cells are read directly, ghosts through fetch_r
template <bool CART, bool GEN>
__device__ __forceinline__ double ew_value(const StageArgs& A, const DevBlock& B, int k, int j, int v,
                                           const Phys& ph) {
    if (!(GEN && ph.prim) && k >= 0 && k < B.ny && j >= 0 && j < B.nx)
        return gld(A.Uin, B.off + 4 * ((long)k * B.nx + j) + v);
    const Q4 q = fetch_r<CART, GEN>(A.blocks, B, A.Uin, k, j, ph);
    return v == 0 ? q.v[0] : (v == 1 ? q.v[1] : (v == 2 ? q.v[2] : q.v[3]));
}

template <bool CART, bool GEN>
__device__ __forceinline__ void ew_recon_tile(const StageArgs& A, const WorkItem it) {
    const DevBlock& B = A.blocks[it.block];
    if (A.skip_dead && A.bad[B.case_id] < *A.step) return;
    const int c = (int)threadIdx.x >> 2, v = (int)threadIdx.x & 3;
    if (c >= it.r0) return;
    const int e = it.j0 + c;
    const int k = e / B.nx, j = e - k * B.nx;
    const Phys ph = A.ph;
    const double u = ew_value<CART, GEN>(A, B, k, j, v, ph);
    const double uw = ew_value<CART, GEN>(A, B, k, j - 1, v, ph);
    const double ue = ew_value<CART, GEN>(A, B, k, j + 1, v, ph);
    const double us = ew_value<CART, GEN>(A, B, k - 1, j, v, ph);
    const double un = ew_value<CART, GEN>(A, B, k + 1, j, v, ph);
    const phx::CellGeo g = cell_geo<CART>(B, k, j);
    double qE, qW, qN, qS;
    phx::reconstruct_var<GEN>(u, uw, ue, us, un, g, ph.limiter, qE, qW, qN, qS);
    double* rb = A.rb + phx::adj::RB * (B.off / 4 + e) + v;
    rb[phx::adj::RB_E] = qE;
    rb[phx::adj::RB_W] = qW;
    rb[phx::adj::RB_N] = qN;
    rb[phx::adj::RB_S] = qS;
}

template <bool CART, int FLUX, bool GEN>
__device__ __forceinline__ void ew_face_tile(const StageArgs& A, const WorkItem it) {
    const DevBlock& B = A.blocks[it.block];
    if (A.skip_dead && A.bad[B.case_id] < *A.step) return;
    if ((int)threadIdx.x >= it.r0) return;
    const long long f = (long long)it.j0 + threadIdx.x;
    const phx::host::AdjBlock O{B.off / 4, B.foff};
    const phx::adj::FaceRef F = phx::adj::face_ref(B, f);
    Q4 L, R;
    phx::adj::face_states<CART>(A.blocks, B, A.Uin, A.rb, O, F, L, R, A.ph);
    const Q4 Fl = F.xdir ? xflux<CART, FLUX, GEN>(B, L, R, F.k, F.j, A.ph)
                         : yflux<CART, FLUX, GEN>(B, L, R, F.k, F.j, A.ph);
    st4(A.ff + 4 * (B.foff + f), Fl);
}

template <bool CART>
__device__ __forceinline__ void ew_update_tile(const StageArgs& A, const WorkItem it) {
    const DevBlock& B = A.blocks[it.block];
    if (A.skip_dead && A.bad[B.case_id] < *A.step) return;  // uniform over the workgroup
    const int t = (int)threadIdx.x;
    double lane_dt = __longlong_as_double((long long)DT_INF_BITS);
    bool lane_bad = false;
    if (t < it.r0) {
        const int e = it.j0 + t;
        const int k = e / B.nx, j = e - k * B.nx;
        const phx::host::AdjBlock O{B.off / 4, B.foff};
        const Q4 Res = phx::adj::cell_residual<CART>(B, O, A.ff, k, j);
        const double dt = (A.dt != nullptr) ? ld_const(A.dt, B.case_id) : 0.0;
        emit_row<CART>(A, B, k, j, dt, Res, lane_dt, lane_bad);
    }
    reduce_item(A, B, t & (WAVE - 1), lane_dt, lane_bad);  // every lane of the wave takes part
}

template <int CK, bool GEN>
__global__ void __launch_bounds__(EW_WG) k_ew_recon(StageArgs A) {
    int lo, hi, first, stride;
    item_range(A.nitems, lo, hi, first, stride);
    for (int i = first; i < hi; i += stride) {
        const WorkItem it = ld_const(A.items, i);
        if (CK == 2) {  // items of both geometry kinds: the path of the item's block
            if (A.blocks[it.block].kind == PHX_GEOM_CARTESIAN) ew_recon_tile<true, GEN>(A, it);
            else ew_recon_tile<false, GEN>(A, it);
        } else {
            ew_recon_tile<CK == 1, GEN>(A, it);
        }
    }
}

template <int CK, int FLUX, bool GEN>
__global__ void __launch_bounds__(EW_WG) k_ew_face(StageArgs A) {
    int lo, hi, first, stride;
    item_range(A.nitems, lo, hi, first, stride);
    for (int i = first; i < hi; i += stride) {
        const WorkItem it = ld_const(A.items, i);
        if (CK == 2) {
            if (A.blocks[it.block].kind == PHX_GEOM_CARTESIAN) ew_face_tile<true, FLUX, GEN>(A, it);
            else ew_face_tile<false, FLUX, GEN>(A, it);
        } else {
            ew_face_tile<CK == 1, FLUX, GEN>(A, it);
        }
    }
}

template <int CK>
__global__ void __launch_bounds__(EW_WG) k_ew_update(StageArgs A) {
    int lo, hi, first, stride;
    item_range(A.nitems, lo, hi, first, stride);
    for (int i = first; i < hi; i += stride) {
        const WorkItem it = ld_const(A.items, i);
        if (CK == 2) {
            if (A.blocks[it.block].kind == PHX_GEOM_CARTESIAN) ew_update_tile<true>(A, it);
            else ew_update_tile<false>(A, it);
        } else {
            ew_update_tile<CK == 1>(A, it);
        }
    }
}

// standalone CFL reduction over the owned cells of every work item
struct DtArgs {
    const DevBlock* blocks;
    const WorkItem* items;
    int nitems;
    int strip;  // owned columns per item of this work list
    const double* U;
    unsigned long long* dtmin;
    int* bad;
    const int* step;
    int nslot;
    Phys ph;
};

template <bool CART>
__global__ void __launch_bounds__(WAVE) k_dt(DtArgs A) {
    int lo, hi, first, stride;
    item_range(A.nitems, lo, hi, first, stride);
    const int lane = threadIdx.x & (WAVE - 1);
    for (int i = first; i < hi; i += stride) {
        const WorkItem it = ld_const(A.items, i);
        const DevBlock& B = A.blocks[it.block];
        const int j = it.j0 + lane;
        const bool own = (lane < A.strip) && (j < B.nx);
        double lane_dt = __longlong_as_double((long long)DT_INF_BITS);
        bool lane_bad = false;
        if (own) {
            for (int k = it.r0; k < it.r1; ++k) {
                const Q4 q = ld4(A.U + B.off + 4 * ((long)k * B.nx + j));
                double dx, dy;
                cell_dxdy<CART>(B, k, j, dx, dy);
                const double t = phx::cell_dt(q, dx, dy, A.ph);
                if (!(q.v[0] > 0.0) || !(q.v[3] > 0.0) || !(t == t)) lane_bad = true;
                else lane_dt = fmin(lane_dt, t);
            }
        }
        double m = lane_dt;
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) m = fmin(m, __shfl_xor(m, o, WAVE));
        const bool anybad = __any(lane_bad);
        if (lane == 0) {
            const unsigned long long bits = (unsigned long long)__double_as_longlong(m);
            if (bits < DT_INF_BITS) atomicMin(&A.dtmin[dt_slot(B.case_id, wave_slot(A.nslot), A.nslot)], bits);
            if (anybad) atomicMin(&A.bad[B.case_id], *A.step);
        }
    }
}

// dt = CFL*min, clipped to t_final - t (solvers/base.py:126-136); advance t (Euler2D.py:209).
// One workgroup; a case that failed at an earlier step gets dt = 0 and keeps its time.
// One lane per slot (nslot lanes per case, one memory round trip), the minima combined by shuffles
// inside the lane group; slots cleared when advancing.  The case loop bound is uniform, so every
// lane takes part in the shuffles.  Launched as one workgroup of FIN_THREADS threads.
constexpr int FIN_THREADS = 1024;
__global__ void __launch_bounds__(FIN_THREADS) k_finalize(int ncases, int nslot, double cfl, unsigned long long* dtmin,
                                                          const double* tfinal, double* t, double* dt, int* step,
                                                          const int* bad, const double* frozen, int advance) {
    const int s_next = step[0] + 1;
    const int g = threadIdx.x & (nslot - 1);
    const int per_pass = blockDim.x / nslot;
    for (int base = 0; base < ncases; base += per_pass) {
        const int c = base + threadIdx.x / nslot;
        const bool active = c < ncases;
        unsigned long long m = DT_INF_BITS;
        if (active && frozen == nullptr) m = dtmin[dt_slot(c, g, nslot)];
        for (int o = nslot / 2; o > 0; o >>= 1) {
            const unsigned long long x = __shfl_xor(m, o, nslot);
            m = x < m ? x : m;
        }
        if (!active) continue;
        if (advance) dtmin[dt_slot(c, g, nslot)] = DT_INF_BITS;
        if (g != 0) continue;
        double d;
        if (frozen != nullptr) {
            d = frozen[c];
        } else {
            d = cfl * __longlong_as_double((long long)m);
            const double rem = tfinal[c] - t[c];
            if (rem < d) d = rem;
        }
        if (advance && bad[c] < s_next) d = 0.0;
        dt[c] = d;
        if (advance) t[c] = t[c] + d;
    }
    __syncthreads();
    if (advance && threadIdx.x == 0) step[0] = s_next;
}
__global__ void k_reset(int ncases, int nslot, const double* t0, double* t, unsigned long long* dtmin, int* bad,
                        int* step, int step0) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;  // one thread per (case, slot)
    if (c < ncases * nslot) dtmin[(long)c * DT_STRIDE] = DT_INF_BITS;
    if (c < ncases) {
        t[c] = t0[c];
        bad[c] = 0x7fffffff;
    }
    if (c == 0) step[0] = step0;
}

__global__ void k_fill_inf(int ncases, int nslot, unsigned long long* dtmin) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // one thread per (case, slot)
    if (i < ncases * nslot) dtmin[(long)i * DT_STRIDE] = DT_INF_BITS;
}

// ghost strips for tests: block b, cell index i over E(ny) W(ny) N(nx) S(nx)
__global__ void k_ghosts(const DevBlock* blocks, int nblocks, const long long* goff, const double* U, double* G) {
    const int b = blockIdx.x;
    if (b >= nblocks) return;
    const DevBlock& B = blocks[b];
    const int n = 2 * B.ny + 2 * B.nx;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int k, j;
        if (i < B.ny) { k = i; j = B.nx; }
        else if (i < 2 * B.ny) { k = i - B.ny; j = -1; }
        else if (i < 2 * B.ny + B.nx) { k = B.ny; j = i - 2 * B.ny; }
        else { k = -1; j = i - 2 * B.ny - B.nx; }
        const Q4 q = (B.kind == PHX_GEOM_CARTESIAN) ? fetch<true>(blocks, B, U, k, j) : fetch<false>(blocks, B, U, k, j);
        st4(G + goff[b] + 4L * i, q);
    }
}

}  // namespace

// =============================================================================================
// host side
// =============================================================================================
namespace {
thread_local std::string g_err;
}  // namespace

int phx::host::fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

namespace {

double* buf_ptr(phx_plan* P, double* U, int id) {
    if (id == BUF_USER) return U;
    if (id == BUF_S0) return P->sbuf[0];
    if (id == BUF_S1) return P->sbuf[1];
    return P->acc[id - BUF_ACC0];
}

// Butcher table -> per-stage program in accumulator form: U^(s+1) = P_s + (dt a_ss) k_s and
// P_t += (dt a_ts) k_s for later rows t, each P_t starting from U0; zero coefficients skipped,
// which reproduces the summation order of ExplicitRungeKutta.integrate (explicit_runge_kutta.py:63-75).
int build_programs(phx_plan* P, int S, const double* a) {
    std::vector<int> Pt(S, BUF_USER);
    std::vector<int> free_acc;
    int nacc = 0;
    P->progs.clear();
    for (int s = 0; s < S; ++s) {
        StageProg g;
        g.in = (s == 0) ? BUF_USER : ((s - 1) % 2 == 0 ? BUF_S0 : BUF_S1);
        // the last stage writes the caller's state in place unless it also reads it as stencil input
        const int out = (s == S - 1 && S > 1) ? BUF_USER : (s % 2 == 0 ? BUF_S0 : BUF_S1);
        g.outs.push_back({out, Pt[s], a[s * S + s]});
        g.state_out = (s == S - 1) ? 0 : -1;
        if (Pt[s] >= BUF_ACC0) free_acc.push_back(Pt[s]);
        for (int t = s + 1; t < S; ++t) {
            const double c = a[t * S + s];
            if (c == 0.0) continue;
            if (Pt[t] == BUF_USER) {
                int id;
                if (!free_acc.empty()) { id = free_acc.back(); free_acc.pop_back(); }
                else { if (nacc >= MAX_ACC) return fail(PHX_E_INVALID, "Butcher table needs too many accumulators"); id = BUF_ACC0 + nacc++; }
                g.outs.push_back({id, BUF_USER, c});
                Pt[t] = id;
            } else {
                g.outs.push_back({Pt[t], Pt[t], c});
            }
        }
        if ((int)g.outs.size() > MAX_OUT) return fail(PHX_E_INVALID, "Butcher stage has too many outputs");
        P->progs.push_back(g);
    }
    P->nacc = nacc;
    P->copy_back = (S == 1);
    return PHX_OK;
}

int upload_items(std::vector<WorkItem> (&items)[2], WorkItem* (&ditems)[2], int (&grid)[2], int grid_waves) {
    for (int kind = 0; kind < 2; ++kind) {
        const size_t n = items[kind].size();
        if (n) {
            PHX_CK(hipMalloc(&ditems[kind], n * sizeof(WorkItem)));
            PHX_CK(hipMemcpy(ditems[kind], items[kind].data(), n * sizeof(WorkItem), hipMemcpyHostToDevice));
        }
        grid[kind] = (int)((grid_waves > 0 && (size_t)grid_waves < n) ? grid_waves : n);
    }
    return PHX_OK;
}

// Work lists: per block, row ranges of `rows` rows times column strips.  The fused variant uses
// STRIP owned columns per item; the split variant uses STRIP_A for its reconstruction kernel,
// STRIP_B with `rows` face rows (ny + 1 of them) for the y-face kernel and STRIP_B with rows_b rows
// for the update kernel.
void release_graphs(phx_plan* P);

// flat y-face stage: lanes per workgroup (rows_b + 1 waves, as the one-face-per-wave stage) and the
// faces its dynamic LDS may hold (64 KiB)
constexpr int FLAT_MAX_FACES = 2048;
int flat_threads(const phx_plan* P) { return (std::min(std::max(P->rows_b, 1), FLAT_MAX_WAVES - 1) + 1) * WAVE; }

// The fused kernel's work list of one geometry kind, arranged for the XCDs.  Workgroup w runs item
// lo_x + (w >> 3) of XCD x = w & 7, whose items are the contiguous range [n x / 8, n (x + 1) / 8)
// (item_range), so each XCD works through its range in list order.
//  * Pairs of consecutive items (neighbouring column strips, or row ranges, of one block) are dealt
//    to the XCDs in turn, so every XCD gets an even share of every case: cases differ in cost per
//    cell (uniform regions skip divisions, cases that failed are skipped), and contiguous ranges of
//    whole cases left some XCDs with less work than others.  (Dealing whole row chunks, all strips
//    of one row range, was 1 % slower on the 64-case batch; single items the same as pairs.)
//  * Every range is made the same length and its last items are split into two halves of their
//    rows: an XCD then finishes on half-size items, which levels the last round of workgroups.
// `tail` = splits per XCD; the set of cells and the work per cell are unchanged.  (Should a range run
// out of items of two or more rows before reaching the common length, the ranges of item_range
// shift against these groups: every item still runs exactly once.)
void order_for_xcds(std::vector<WorkItem>& L, int tail) {
    const int nb = (int)L.size();
    if (tail <= 0 || nb < 8 * 4) return;
    tail = std::min(tail, nb / 8 / 2);
    std::vector<WorkItem> per[8];
    for (int i = 0, g = 0; i < nb; ++g) {
        const int k = std::min(nb, i + 2);
        per[g % 8].insert(per[g % 8].end(), L.begin() + i, L.begin() + k);
        i = k;
    }
    size_t most = 0;
    for (int x = 0; x < 8; ++x) most = std::max(most, per[x].size());
    const int T = (int)most + tail;  // items per XCD after the splits
    std::vector<WorkItem> out;
    out.reserve((size_t)8 * T);
    for (int x = 0; x < 8; ++x) {
        int need = T - (int)per[x].size();
        std::vector<WorkItem>& head = per[x];
        std::vector<WorkItem> back;
        while (need > 0 && !head.empty()) {
            const WorkItem it = head.back();
            head.pop_back();
            if (it.r1 - it.r0 < 2) {
                back.push_back(it);
                continue;
            }
            const int m = it.r0 + (it.r1 - it.r0) / 2;
            back.push_back(WorkItem{it.block, it.j0, m, it.r1});
            back.push_back(WorkItem{it.block, it.j0, it.r0, m});
            --need;
        }
        out.insert(out.end(), head.begin(), head.end());
        out.insert(out.end(), back.rbegin(), back.rend());
    }
    L.swap(out);
}

// A work list dealt to the XCDs in chunks of `chunk` consecutive items: chunk c goes to XCD c mod 8
// (the next XCD with room once one is full), each XCD keeping the range length item_range gives it,
// so every XCD gets a share of every region of the mesh.  Chosen per workload by the autotuner:
// it pays where the cost per cell varies over the mesh (uniform regions skip divisions), and costs
// L2 reuse of the items' shared halo rows and columns where it does not.
void deal_to_xcds(std::vector<WorkItem>& L, int chunk) {
    const int n = (int)L.size();
    if (chunk <= 0 || n < 8 * chunk) return;
    int lo[8], room[8];
    for (int x = 0; x < 8; ++x) {
        lo[x] = (int)((long)n * x / 8);
        room[x] = (int)((long)n * (x + 1) / 8) - lo[x];
    }
    std::vector<WorkItem> out(L.size());
    int used[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0, x = 0; i < n; i += chunk, x = (x + 1) % 8)
        for (int q = i; q < std::min(n, i + chunk); ++q) {
            int t = x;
            while (used[t] >= room[t]) t = (t + 1) % 8;
            out[lo[t] + used[t]++] = L[q];
        }
    L.swap(out);
}

int build_items(phx_plan* P) {
    for (int kind = 0; kind < 2; ++kind) {
        P->items[kind].clear();
        P->items_b[kind].clear();
        P->items_f[kind].clear();
        if (P->ditems[kind]) { (void)hipFree(P->ditems[kind]); P->ditems[kind] = nullptr; }
        if (P->ditems_b[kind]) { (void)hipFree(P->ditems_b[kind]); P->ditems_b[kind] = nullptr; }
        if (P->ditems_f[kind]) { (void)hipFree(P->ditems_f[kind]); P->ditems_f[kind] = nullptr; }
        for (int p = 0; p < 3; ++p) {
            P->items_e[p][kind].clear();
            if (P->ditems_e[p][kind]) { (void)hipFree(P->ditems_e[p][kind]); P->ditems_e[p][kind] = nullptr; }
        }
    }
    for (int p = 0; p < 3; ++p) {
        if (P->ditems_em[p]) { (void)hipFree(P->ditems_em[p]); P->ditems_em[p] = nullptr; }
        P->n_em[p] = P->grid_em[p] = 0;
        if (P->ditems_m[p]) { (void)hipFree(P->ditems_m[p]); P->ditems_m[p] = nullptr; }
        P->n_m[p] = P->grid_m[p] = 0;
    }
    const int rows = P->rows_per_item > 0 ? P->rows_per_item : 32;
    const bool split = P->variant == 1 || P->variant >= 5;
    P->flat_q[0] = P->flat_q[1] = 0;
    const int strip = split ? STRIP_A : STRIP;
    const int item_rows = P->variant == 2 ? WGW * rows : rows;  // shared-halo items: WGW waves of `rows`
    for (int b = 0; b < P->nblocks; ++b) {
        const phx_block_desc& d = P->desc[b];
        const int kind = d.geom_kind == PHX_GEOM_CARTESIAN ? 1 : 0;
        for (int r0 = 0; r0 < d.ny; r0 += item_rows)
            for (int j0 = 0; j0 < d.nx; j0 += strip)
                P->items[kind].push_back({b, j0, r0, std::min(r0 + item_rows, d.ny)});
        if (P->variant == 3) {
            const int nc = d.nx * d.ny, nf = d.ny * (d.nx + 1) + (d.ny + 1) * d.nx;
            for (int e = 0; e < nc; e += EW_RCELLS) P->items_e[0][kind].push_back({b, e, std::min(EW_RCELLS, nc - e), 0});
            for (int e = 0; e < nf; e += EW_WG) P->items_e[1][kind].push_back({b, e, std::min(EW_WG, nf - e), 0});
            for (int e = 0; e < nc; e += EW_WG) P->items_e[2][kind].push_back({b, e, std::min(EW_WG, nc - e), 0});
        }
        if (P->variant == 7 || P->variant == 8)  // row-major reconstruction items: LIN_OWN positions each
            for (int p0 = 0; p0 < d.ny * (d.nx + 1); p0 += LIN_OWN) P->items_f[kind].push_back({b, p0, 0, 0});
        if (P->variant == 8) {
            // flat y-face ranges: workgroups of T lanes, q >= rows_per_item faces per lane and at
            // least 64 cells per range (n = q*T - nx)
            const int T = flat_threads(P), nc = d.nx * d.ny;
            const int qmin = (d.nx + 64 + T - 1) / T;  // phx_set_variant checked qmin * T <= FLAT_MAX_FACES
            const int q = std::min(std::max(P->rows_per_item, qmin), FLAT_MAX_FACES / T);
            P->flat_q[kind] = std::max(P->flat_q[kind], q);
            const int n = q * T - d.nx;
            for (int c0 = 0; c0 < nc; c0 += n) P->items_b[kind].push_back({b, c0, std::min(n, nc - c0), 0});
        }
        if (P->variant == 5 || P->variant == 6 || P->variant == 7) {
            const int rb = P->variant != 5 ? std::min(std::max(P->rows_b, 1), YU_MAX_ROWS) : P->rows_b;
            for (int r0 = 0; r0 < d.ny; r0 += rb)
                for (int j0 = 0; j0 < d.nx; j0 += STRIP_B)
                    P->items_b[kind].push_back({b, j0, r0, std::min(r0 + rb, d.ny)});
        }
        if (P->variant == 1) {
            for (int r0 = 0; r0 <= d.ny; r0 += rows)
                for (int j0 = 0; j0 < d.nx; j0 += STRIP_B)
                    P->items_f[kind].push_back({b, j0, r0, std::min(r0 + rows, d.ny + 1)});
            for (int r0 = 0; r0 < d.ny; r0 += P->rows_b)
                for (int j0 = 0; j0 < d.nx; j0 += STRIP_B)
                    P->items_b[kind].push_back({b, j0, r0, std::min(r0 + P->rows_b, d.ny)});
        }
    }
    // fused kernel, one workgroup per item: item pairs dealt to the XCDs, half a round of
    // workgroups per XCD (3 waves on each of its SIMDs) ending on half-size items
    if (P->variant == 0 && P->grid_waves == 0)
        for (int kind = 0; kind < 2; ++kind) order_for_xcds(P->items[kind], P->ncu * 4 * 3 / 8 / 2);
    if (P->variant != 0 && P->grid_waves == 0 && P->xcd_chunk > 0)
        for (int kind = 0; kind < 2; ++kind) {
            deal_to_xcds(P->items[kind], P->xcd_chunk);
            deal_to_xcds(P->items_b[kind], P->xcd_chunk);
            deal_to_xcds(P->items_f[kind], P->xcd_chunk);
            for (int p = 0; p < 3; ++p) deal_to_xcds(P->items_e[p][kind], P->xcd_chunk);
        }
    int rc = upload_items(P->items, P->ditems, P->grid, P->grid_waves);
    if (rc) return rc;
    rc = upload_items(P->items_b, P->ditems_b, P->grid_b, P->grid_waves);
    if (rc) return rc;
    rc = upload_items(P->items_f, P->ditems_f, P->grid_f, P->grid_waves);
    if (rc) return rc;
    for (int p = 0; p < 3; ++p) {
        rc = upload_items(P->items_e[p], P->ditems_e[p], P->grid_e[p], P->grid_waves);
        if (rc) return rc;
    }
    // both geometry kinds: one list per phase of the element-wise stage and of the split stages 6
    // and 7 with midpoint faces (one launch instead of two; the kernels choose the path per block)
    auto merge = [&](const std::vector<WorkItem> (&two)[2], WorkItem*& dev, int& n_out, int& grid_out) -> int {
        if (two[0].empty() || two[1].empty()) return PHX_OK;
        std::vector<WorkItem> both(two[0]);
        both.insert(both.end(), two[1].begin(), two[1].end());
        if (P->grid_waves == 0) deal_to_xcds(both, P->xcd_chunk);
        const size_t n = both.size();
        PHX_CK(hipMalloc(&dev, n * sizeof(WorkItem)));
        PHX_CK(hipMemcpy(dev, both.data(), n * sizeof(WorkItem), hipMemcpyHostToDevice));
        n_out = (int)n;
        grid_out = (int)((P->grid_waves > 0 && (size_t)P->grid_waves < n) ? P->grid_waves : n);
        return PHX_OK;
    };
    if (P->variant == 3)
        for (int p = 0; p < 3; ++p) {
            rc = merge(P->items_e[p], P->ditems_em[p], P->n_em[p], P->grid_em[p]);
            if (rc) return rc;
        }
    if ((P->variant == 6 || P->variant == 7) && P->ph.nq == 1) {
        rc = merge(P->items_b, P->ditems_m[1], P->n_m[1], P->grid_m[1]);
        if (rc) return rc;
        if (P->variant == 6) rc = merge(P->items, P->ditems_m[0], P->n_m[0], P->grid_m[0]);
        else rc = merge(P->items_f, P->ditems_m[2], P->n_m[2], P->grid_m[2]);
        if (rc) return rc;
    }
    release_graphs(P);
    return PHX_OK;
}

// Destroys the captured step graphs (they hold the launch list and the state pointer).
void release_graphs(phx_plan* P) {
    for (int g = 0; g < 2; ++g) {
        if (P->gexec[g]) { (void)hipGraphExecDestroy(P->gexec[g]); P->gexec[g] = nullptr; }
        if (P->graph[g]) { (void)hipGraphDestroy(P->graph[g]); P->graph[g] = nullptr; }
    }
    P->graph_U = nullptr;
}

StageArgs base_args(phx_plan* P) {
    StageArgs A;
    memset(&A, 0, sizeof(A));
    A.blocks = P->dblocks;
    A.state_out = -1;
    A.dt = P->dt;
    A.bad = P->bad;
    A.nslot = P->nslot;
    A.step = P->step;
    A.ph = P->ph;
    return A;
}

// Launchers of the stage kernel families.  M = requested waves per SIMD, G = generic limiter.
// The instantiated register budgets are limited to those that compile without heavy spilling:
// fused <= 3 (2 for the generic limiter), reconstruction 3-4 (3 generic); the y-face and update
// kernels need no limiter and are built once.
struct k_stage_tag {
    template <bool C, int F, int M, bool G>
    static void launch(dim3 g, hipStream_t st, const StageArgs& A) {
        constexpr int MM = G ? 2 : (M > 3 ? 3 : M);
        hipLaunchKernelGGL((k_stage<C, F, MM, G>), g, dim3(WAVE), 0, st, A);
    }
};
struct k_stage_wg_tag {
    template <bool C, int F, int M, bool G>
    static void launch(dim3 g, hipStream_t st, const StageArgs& A) {
        hipLaunchKernelGGL((k_stage_wg<C, F, 2, G>), g, dim3(WGW * WAVE), 0, st, A);
    }
};
struct k_stage_pair_tag {
    template <bool C, int F, int M, bool G>
    static void launch(dim3 g, hipStream_t st, const StageArgs& A) {
        constexpr int MM = G ? 2 : (M > 3 ? 3 : (M < 2 ? 2 : M));
        hipLaunchKernelGGL((k_stage_pair<C, F, MM, G>), g, dim3(PAIR_WAVES * WAVE), 0, st, A);
    }
};
struct k_recon_x_tag {
    template <bool C, int F, int M, bool G>
    static void launch(dim3 g, hipStream_t st, const StageArgs& A) {
        constexpr int MM = G ? 3 : (M < 3 ? 3 : M);
        hipLaunchKernelGGL((k_recon_x<C ? 1 : 0, F, MM, G>), g, dim3(WAVE), 0, st, A);
    }
};
struct k_recon_x_lin_tag {
    template <bool C, int F, int M, bool G>
    static void launch(dim3 g, hipStream_t st, const StageArgs& A) {
        constexpr int MM = G ? 3 : (M < 3 ? 3 : M);
        hipLaunchKernelGGL((k_recon_x_lin<C ? 1 : 0, F, MM, G>), g, dim3(WAVE), 0, st, A);
    }
};
struct k_yface_tag {
    template <bool C, int F, int M, bool G>
    static void launch(dim3 g, hipStream_t st, const StageArgs& A) {
        hipLaunchKernelGGL((k_yface<C, F, G>), g, dim3(WAVE), 0, st, A);
    }
};
struct k_recon_x_q_tag {
    template <bool C, int F, int M, bool G>
    static void launch(dim3 g, hipStream_t st, const StageArgs& A) {
        hipLaunchKernelGGL((k_recon_x_q<C, F>), g, dim3(WAVE), 0, st, A);
    }
};
struct k_yface_q_tag {
    template <bool C, int F, int M, bool G>
    static void launch(dim3 g, hipStream_t st, const StageArgs& A) {
        hipLaunchKernelGGL((k_yface_q<C, F>), g, dim3(WAVE), 0, st, A);
    }
};
struct k_ew_recon_tag {
    template <bool C, int F, int M, bool G>
    static void launch(dim3 g, hipStream_t st, const StageArgs& A) {
        hipLaunchKernelGGL((k_ew_recon<C ? 1 : 0, G>), g, dim3(EW_WG), 0, st, A);
    }
};
struct k_ew_face_tag {
    template <bool C, int F, int M, bool G>
    static void launch(dim3 g, hipStream_t st, const StageArgs& A) {
        hipLaunchKernelGGL((k_ew_face<C ? 1 : 0, F, G>), g, dim3(EW_WG), 0, st, A);
    }
};
struct k_ew_update_tag {
    template <bool C, int F, int M, bool G>
    static void launch(dim3 g, hipStream_t st, const StageArgs& A) {
        hipLaunchKernelGGL((k_ew_update<C ? 1 : 0>), g, dim3(EW_WG), 0, st, A);
    }
};
struct k_yface_update_tag {
    template <bool C, int F, int M, bool G>
    static void launch(dim3 g, hipStream_t st, const StageArgs& A) {
        hipLaunchKernelGGL((k_yface_update<C, F, G>), g, dim3(WAVE), 0, st, A);
    }
};
struct k_yface_update_q_tag {
    template <bool C, int F, int M, bool G>
    static void launch(dim3 g, hipStream_t st, const StageArgs& A) {
        hipLaunchKernelGGL((k_yface_update_q<C, F>), g, dim3(WAVE), 0, st, A);
    }
};
struct k_yface_update_wg_tag {
    template <bool C, int F, int M, bool G>
    static void launch(dim3 g, hipStream_t st, const StageArgs& A) {
        hipLaunchKernelGGL((k_yface_update_wg<C ? 1 : 0, F, G, false>), g, dim3((A.rows_wave + 1) * WAVE), 0, st, A);
    }
};
struct k_yface_update_wg_q_tag {
    template <bool C, int F, int M, bool G>
    static void launch(dim3 g, hipStream_t st, const StageArgs& A) {
        hipLaunchKernelGGL((k_yface_update_wg<C ? 1 : 0, F, true, true>), g, dim3((A.rows_wave + 1) * WAVE), 0, st, A);
    }
};
struct k_yface_update_flat_tag {
    template <bool C, int F, int M, bool G>
    static void launch(dim3 g, hipStream_t st, const StageArgs& A) {
        const int T = (A.rows_wave + 1) * WAVE;
        hipLaunchKernelGGL((k_yface_update_flat<C, F, G>), g, dim3(T), (size_t)A.fq * T * sizeof(Q4), st, A);
    }
};
struct k_update_tag {
    template <bool C, int F, int M, bool G>
    static void launch(dim3 g, hipStream_t st, const StageArgs& A) {
        hipLaunchKernelGGL((k_update<C>), g, dim3(WAVE), 0, st, A);
    }
};

// the generic instantiation: another limiter, primitive-variable reconstruction or Gauss faces
bool generic_instance(const phx_plan* P) {
    return P->ph.limiter != PHX_LIMITER_VENKATAKRISHNAN || P->ph.prim || P->ph.nq > 1;
}

// one launch of phase p (0 reconstruction, 1 faces, 2 update) of the element-wise stage over the
// merged list of both geometry kinds
void launch_ew_mixed(const phx_plan* P, StageArgs A, int p, hipStream_t st) {
    A.items = P->ditems_em[p];
    A.nitems = P->n_em[p];
    const dim3 g(P->grid_em[p]), b(EW_WG);
    const bool gen = generic_instance(P), roe = P->flux == PHX_FLUX_ROE;
    if (p == 0) {
        if (gen) hipLaunchKernelGGL((k_ew_recon<2, true>), g, b, 0, st, A);
        else hipLaunchKernelGGL((k_ew_recon<2, false>), g, b, 0, st, A);
    } else if (p == 1) {
        if (roe && gen) hipLaunchKernelGGL((k_ew_face<2, 0, true>), g, b, 0, st, A);
        else if (roe) hipLaunchKernelGGL((k_ew_face<2, 0, false>), g, b, 0, st, A);
        else if (gen) hipLaunchKernelGGL((k_ew_face<2, 1, true>), g, b, 0, st, A);
        else hipLaunchKernelGGL((k_ew_face<2, 1, false>), g, b, 0, st, A);
    } else {
        hipLaunchKernelGGL((k_ew_update<2>), g, b, 0, st, A);
    }
}

// one launch of the reconstruction kernel of variant 6 (lin = false) or 7 (lin = true) over the
// merged list of both geometry kinds
void launch_recon_mixed(const phx_plan* P, StageArgs A, bool lin, hipStream_t st) {
    const int L = lin ? 2 : 0;
    A.items = P->ditems_m[L];
    A.nitems = P->n_m[L];
    const dim3 g(P->grid_m[L]), b(WAVE);
    const bool gen = generic_instance(P), roe = P->flux == PHX_FLUX_ROE;
    const bool w4 = !gen && P->waves_per_simd >= 4;  // the register budgets the per-kind launches use
#define PHX_RX(K, F, M, G) hipLaunchKernelGGL((K<2, F, M, G>), g, b, 0, st, A)
    if (lin) {
        if (gen) { if (roe) PHX_RX(k_recon_x_lin, 0, 3, true); else PHX_RX(k_recon_x_lin, 1, 3, true); }
        else if (w4) { if (roe) PHX_RX(k_recon_x_lin, 0, 4, false); else PHX_RX(k_recon_x_lin, 1, 4, false); }
        else { if (roe) PHX_RX(k_recon_x_lin, 0, 3, false); else PHX_RX(k_recon_x_lin, 1, 3, false); }
    } else {
        if (gen) { if (roe) PHX_RX(k_recon_x, 0, 3, true); else PHX_RX(k_recon_x, 1, 3, true); }
        else if (w4) { if (roe) PHX_RX(k_recon_x, 0, 4, false); else PHX_RX(k_recon_x, 1, 4, false); }
        else { if (roe) PHX_RX(k_recon_x, 0, 3, false); else PHX_RX(k_recon_x, 1, 3, false); }
    }
#undef PHX_RX
}

// one launch of the y-face and update kernel of variants 6 and 7 over the merged list
void launch_yface_wg_mixed(const phx_plan* P, StageArgs A, hipStream_t st) {
    A.items = P->ditems_m[1];
    A.nitems = P->n_m[1];
    const dim3 g(P->grid_m[1]), b((A.rows_wave + 1) * WAVE);
    const bool gen = generic_instance(P), roe = P->flux == PHX_FLUX_ROE;
    if (roe && gen) hipLaunchKernelGGL((k_yface_update_wg<2, 0, true, false>), g, b, 0, st, A);
    else if (roe) hipLaunchKernelGGL((k_yface_update_wg<2, 0, false, false>), g, b, 0, st, A);
    else if (gen) hipLaunchKernelGGL((k_yface_update_wg<2, 1, true, false>), g, b, 0, st, A);
    else hipLaunchKernelGGL((k_yface_update_wg<2, 1, false, false>), g, b, 0, st, A);
}

// one launch of kernel family T over the items of one geometry kind; minw = register budget in
// waves per SIMD (2, 3 or 4)
template <class T>
void launch_kind(const phx_plan* P, StageArgs A, int kind, const WorkItem* items, int nitems, int grid, int minw,
                 hipStream_t st) {
    A.items = items;
    A.nitems = nitems;
    const dim3 g(grid);
    const bool roe = P->flux == PHX_FLUX_ROE;
    const bool gen = generic_instance(P);
#define PHX_LAUNCH_M(C, F)                                                                       \
    do {                                                                                         \
        if (gen) T::template launch<C, F, 2, true>(g, st, A);                                    \
        else if (minw == 4) T::template launch<C, F, 4, false>(g, st, A);                        \
        else if (minw == 3) T::template launch<C, F, 3, false>(g, st, A);                        \
        else T::template launch<C, F, 2, false>(g, st, A);                                       \
    } while (0)
    if (kind == 1) {
        if (roe) PHX_LAUNCH_M(true, 0);
        else PHX_LAUNCH_M(true, 1);
    } else {
        if (roe) PHX_LAUNCH_M(false, 0);
        else PHX_LAUNCH_M(false, 1);
    }
#undef PHX_LAUNCH_M
}

// One phase of a stage over both geometry kinds: launch(kind, stream) for each kind that has work.
// With work of both kinds, kind 1 runs on the plan's side stream concurrently with kind 0 on st
// (fork and join through events, which a graph capture turns into parallel branches): the kinds
// write disjoint blocks and read only what earlier phases wrote (inter-block outer states are cell
// averages of the stage input), and their per-case CFL and realizability reductions are atomic
// minima, so the order of the two launches does not change any result.  On mixed meshes of a few
// thousand cells the launches, not the arithmetic, set the time of a phase.
template <class F>
void for_kinds(phx_plan* P, bool has0, bool has1, hipStream_t st, F&& launch) {
    if (has0 && has1 && P->side_stream && hipEventRecord(P->ev_fork, st) == hipSuccess &&
        hipStreamWaitEvent(P->side_stream, P->ev_fork, 0) == hipSuccess) {
        launch(1, P->side_stream);
        launch(0, st);
        if (hipEventRecord(P->ev_join, P->side_stream) != hipSuccess ||
            hipStreamWaitEvent(st, P->ev_join, 0) != hipSuccess)
            (void)hipStreamSynchronize(P->side_stream);  // join on the host instead
        return;
    }
    if (has0) launch(0, st);
    if (has1) launch(1, st);
}

void launch_stage(phx_plan* P, StageArgs A, hipStream_t st) {
    if (P->variant == 8) {
        A.sn = P->sn;
        A.p1 = P->p1;
        A.rows_wave = flat_threads(P) / WAVE - 1;
        for_kinds(P, !P->items_f[0].empty(), !P->items_f[1].empty(), st, [&](int kind, hipStream_t s) {
            launch_kind<k_recon_x_lin_tag>(P, A, kind, P->ditems_f[kind], (int)P->items_f[kind].size(),
                                           P->grid_f[kind], P->waves_per_simd, s);
        });
        for_kinds(P, !P->items_b[0].empty(), !P->items_b[1].empty(), st, [&](int kind, hipStream_t s) {
            StageArgs Ak = A;
            Ak.fq = P->flat_q[kind];
            launch_kind<k_yface_update_flat_tag>(P, Ak, kind, P->ditems_b[kind], (int)P->items_b[kind].size(),
                                                 P->grid_b[kind], 2, s);
        });
        return;
    }
    if (P->variant == 7) {
        A.sn = P->sn;
        A.p1 = P->p1;
        A.rows_wave = std::min(std::max(P->rows_b, 1), YU_MAX_ROWS);
        if (P->ditems_m[2]) launch_recon_mixed(P, A, true, st);
        else
            for_kinds(P, !P->items_f[0].empty(), !P->items_f[1].empty(), st, [&](int kind, hipStream_t s) {
                launch_kind<k_recon_x_lin_tag>(P, A, kind, P->ditems_f[kind], (int)P->items_f[kind].size(),
                                               P->grid_f[kind], P->waves_per_simd, s);
            });
        if (P->ditems_m[1]) launch_yface_wg_mixed(P, A, st);
        else
            for_kinds(P, !P->items_b[0].empty(), !P->items_b[1].empty(), st, [&](int kind, hipStream_t s) {
                launch_kind<k_yface_update_wg_tag>(P, A, kind, P->ditems_b[kind], (int)P->items_b[kind].size(),
                                                   P->grid_b[kind], 2, s);
            });
        return;
    }
    if (P->variant == 6) {
        A.sn = P->sn;
        A.p1 = P->p1;
        A.rows_wave = std::min(std::max(P->rows_b, 1), YU_MAX_ROWS);
        const bool gauss = P->ph.nq > 1;
        if (!gauss && P->ditems_m[0] && P->ditems_m[1]) {
            launch_recon_mixed(P, A, false, st);
            launch_yface_wg_mixed(P, A, st);
            return;
        }
        for_kinds(P, !P->items[0].empty(), !P->items[1].empty(), st, [&](int kind, hipStream_t s) {
            if (gauss)
                launch_kind<k_recon_x_q_tag>(P, A, kind, P->ditems[kind], (int)P->items[kind].size(),
                                             P->grid[kind], 2, s);
            else
                launch_kind<k_recon_x_tag>(P, A, kind, P->ditems[kind], (int)P->items[kind].size(),
                                           P->grid[kind], P->waves_per_simd, s);
        });
        for_kinds(P, !P->items_b[0].empty(), !P->items_b[1].empty(), st, [&](int kind, hipStream_t s) {
            if (gauss)
                launch_kind<k_yface_update_wg_q_tag>(P, A, kind, P->ditems_b[kind], (int)P->items_b[kind].size(),
                                                     P->grid_b[kind], 2, s);
            else
                launch_kind<k_yface_update_wg_tag>(P, A, kind, P->ditems_b[kind], (int)P->items_b[kind].size(),
                                                   P->grid_b[kind], 2, s);
        });
        return;
    }
    if (P->variant == 5) {
        A.sn = P->sn;
        A.p1 = P->p1;
        const bool gauss = P->ph.nq > 1;
        for_kinds(P, !P->items[0].empty(), !P->items[1].empty(), st, [&](int kind, hipStream_t s) {
            if (gauss)
                launch_kind<k_recon_x_q_tag>(P, A, kind, P->ditems[kind], (int)P->items[kind].size(),
                                             P->grid[kind], 2, s);
            else
                launch_kind<k_recon_x_tag>(P, A, kind, P->ditems[kind], (int)P->items[kind].size(),
                                           P->grid[kind], P->waves_per_simd, s);
        });
        for_kinds(P, !P->items_b[0].empty(), !P->items_b[1].empty(), st, [&](int kind, hipStream_t s) {
            if (gauss)
                launch_kind<k_yface_update_q_tag>(P, A, kind, P->ditems_b[kind], (int)P->items_b[kind].size(),
                                                  P->grid_b[kind], 2, s);
            else
                launch_kind<k_yface_update_tag>(P, A, kind, P->ditems_b[kind], (int)P->items_b[kind].size(),
                                                P->grid_b[kind], 2, s);
        });
        return;
    }
    if (P->variant == 1) {
        A.sn = P->sn;
        A.p1 = P->p1;
        const bool gauss = P->ph.nq > 1;
        for_kinds(P, !P->items[0].empty(), !P->items[1].empty(), st, [&](int kind, hipStream_t s) {
            if (gauss)
                launch_kind<k_recon_x_q_tag>(P, A, kind, P->ditems[kind], (int)P->items[kind].size(),
                                             P->grid[kind], 2, s);
            else
                launch_kind<k_recon_x_tag>(P, A, kind, P->ditems[kind], (int)P->items[kind].size(),
                                           P->grid[kind], P->waves_per_simd, s);
        });
        A.yf = P->yf;
        for_kinds(P, !P->items_f[0].empty(), !P->items_f[1].empty(), st, [&](int kind, hipStream_t s) {
            if (gauss)
                launch_kind<k_yface_q_tag>(P, A, kind, P->ditems_f[kind], (int)P->items_f[kind].size(),
                                           P->grid_f[kind], 2, s);
            else
                launch_kind<k_yface_tag>(P, A, kind, P->ditems_f[kind], (int)P->items_f[kind].size(),
                                         P->grid_f[kind], 2, s);
        });
        for_kinds(P, !P->items_b[0].empty(), !P->items_b[1].empty(), st, [&](int kind, hipStream_t s) {
            launch_kind<k_update_tag>(P, A, kind, P->ditems_b[kind], (int)P->items_b[kind].size(), P->grid_b[kind],
                                      2, s);
        });
        return;
    }
    if (P->variant == 3) {
        A.rb = P->ew_rb;
        A.ff = P->ew_ff;
        if (P->ditems_em[0]) launch_ew_mixed(P, A, 0, st);
        else
            for_kinds(P, !P->items_e[0][0].empty(), !P->items_e[0][1].empty(), st, [&](int kind, hipStream_t s) {
                launch_kind<k_ew_recon_tag>(P, A, kind, P->ditems_e[0][kind], (int)P->items_e[0][kind].size(),
                                            P->grid_e[0][kind], 2, s);
            });
        if (P->ditems_em[1]) launch_ew_mixed(P, A, 1, st);
        else
            for_kinds(P, !P->items_e[1][0].empty(), !P->items_e[1][1].empty(), st, [&](int kind, hipStream_t s) {
                launch_kind<k_ew_face_tag>(P, A, kind, P->ditems_e[1][kind], (int)P->items_e[1][kind].size(),
                                           P->grid_e[1][kind], 2, s);
            });
        if (P->ditems_em[2]) launch_ew_mixed(P, A, 2, st);
        else
            for_kinds(P, !P->items_e[2][0].empty(), !P->items_e[2][1].empty(), st, [&](int kind, hipStream_t s) {
                launch_kind<k_ew_update_tag>(P, A, kind, P->ditems_e[2][kind], (int)P->items_e[2][kind].size(),
                                             P->grid_e[2][kind], 2, s);
            });
        return;
    }
    if (P->variant == 4) {
        for_kinds(P, !P->items[0].empty(), !P->items[1].empty(), st, [&](int kind, hipStream_t s) {
            launch_kind<k_stage_pair_tag>(P, A, kind, P->ditems[kind], (int)P->items[kind].size(), P->grid[kind],
                                          std::min(P->waves_per_simd, 3), s);
        });
        return;
    }
    if (P->variant == 2) {
        A.rows_wave = P->rows_per_item > 0 ? P->rows_per_item : 1;
        for_kinds(P, !P->items[0].empty(), !P->items[1].empty(), st, [&](int kind, hipStream_t s) {
            launch_kind<k_stage_wg_tag>(P, A, kind, P->ditems[kind], (int)P->items[kind].size(), P->grid[kind],
                                        2, s);
        });
        return;
    }
    // the fused kernel needs more than 128 registers: at most 3 waves per SIMD
    for_kinds(P, !P->items[0].empty(), !P->items[1].empty(), st, [&](int kind, hipStream_t s) {
        launch_kind<k_stage_tag>(P, A, kind, P->ditems[kind], (int)P->items[kind].size(), P->grid[kind],
                                 std::min(P->waves_per_simd, 3), s);
    });
}

void launch_dt_reduce(phx_plan* P, const double* U, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_inf, dim3((P->ncases * P->nslot + 255) / 256), dim3(256), 0, st, P->ncases, P->nslot,
                       P->dtmin);
    for (int kind = 0; kind < 2; ++kind) {
        const int n = (int)P->items[kind].size();
        if (!n) continue;
        DtArgs D{P->dblocks, P->ditems[kind], n, (P->variant == 1 || P->variant >= 5) ? STRIP_A : STRIP,
                 U, P->dtmin,
                 P->bad, P->step, P->nslot,
                 P->ph};
        if (kind == 1) hipLaunchKernelGGL(k_dt<true>, dim3(P->grid[kind]), dim3(WAVE), 0, st, D);
        else hipLaunchKernelGGL(k_dt<false>, dim3(P->grid[kind]), dim3(WAVE), 0, st, D);
    }
}

// Buffers of one stage-program launch: the caller's state U (read as U0 / stage-0 input, written by
// the last stage) and the stage-state buffer the stage reads (s_read) or writes (s_write).
struct BufMap {
    double *u_read, *u_write, *s_read, *s_write;
};

double* map_buf(phx_plan* P, const BufMap& M, int id, bool write) {
    if (id == BUF_USER) return write ? M.u_write : M.u_read;
    if (id == BUF_S0 || id == BUF_S1) return write ? M.s_write : M.s_read;
    return P->acc[id - BUF_ACC0];
}

void launch_prog_stage_mapped(phx_plan* P, size_t s, const BufMap& M, hipStream_t st) {
    const StageProg& g = P->progs[s];
    StageArgs A = base_args(P);
    A.Uin = map_buf(P, M, g.in, false);
    A.nout = (int)g.outs.size();
    for (int m = 0; m < A.nout; ++m) {
        A.dst[m] = map_buf(P, M, g.outs[m].dst, true);
        A.src[m] = map_buf(P, M, g.outs[m].src, false);
        A.a[m] = g.outs[m].a;
    }
    A.state_out = g.state_out;
    A.skip_dead = 1;
    if (g.state_out >= 0) A.dtmin = P->dtmin;
    launch_stage(P, A, st);
}

void launch_prog_stage(phx_plan* P, double* U, size_t s, hipStream_t st) {
    const StageProg& g = P->progs[s];
    int sw = BUF_S0;
    for (const StageOut& o : g.outs)
        if (o.dst == BUF_S0 || o.dst == BUF_S1) sw = o.dst;
    const BufMap M{U, U, buf_ptr(P, U, g.in == BUF_USER ? BUF_S0 : g.in), buf_ptr(P, U, sw)};
    launch_prog_stage_mapped(P, s, M, st);
}

void launch_step(phx_plan* P, double* U, hipStream_t st, const double* frozen) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN_THREADS), 0, st, P->ncases, P->nslot, P->ph.cfl, P->dtmin, P->tfinal, P->t,
                       P->dt, P->step, P->bad, frozen, 1);
    for (size_t s = 0; s < P->progs.size(); ++s) launch_prog_stage(P, U, s, st);
    if (P->copy_back)
        (void)hipMemcpyAsync(U, P->sbuf[0], (size_t)P->state_len * sizeof(double), hipMemcpyDeviceToDevice, st);
}

}  // namespace

extern "C" {

const char* phx_last_error(void) { return g_err.c_str(); }

int64_t phx_state_len(const phx_plan* P) { return P ? P->state_len : -1; }

int phx_plan_destroy(phx_plan* P) {
    if (!P) return PHX_OK;
    (void)hipDeviceSynchronize();
    phx::host::adjoint_release(P);
    release_graphs(P);
    if (P->cap_stream) (void)hipStreamDestroy(P->cap_stream);
    if (P->ev_a) (void)hipEventDestroy(P->ev_a);
    if (P->ev_b) (void)hipEventDestroy(P->ev_b);
    if (P->side_stream) (void)hipStreamDestroy(P->side_stream);
    if (P->ev_fork) (void)hipEventDestroy(P->ev_fork);
    if (P->ev_join) (void)hipEventDestroy(P->ev_join);
    void* ptrs[] = {P->dblocks, P->dgeom, P->ditems[0], P->ditems[1], P->ditems_b[0], P->ditems_b[1], P->ditems_f[0],
                    P->ditems_f[1], P->sn, P->p1, P->yf, P->ew_rb, P->ew_ff, P->ditems_e[0][0], P->ditems_e[0][1],
                    P->ditems_e[1][0], P->ditems_e[1][1], P->ditems_e[2][0], P->ditems_e[2][1],
                    P->ditems_em[0], P->ditems_em[1], P->ditems_em[2], P->ditems_m[0], P->ditems_m[1],
                    P->ditems_m[2],
                    P->sbuf[0], P->sbuf[1], P->dt, P->t, P->tfinal, P->dtmin, P->step, P->bad, P->dghost_off};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (int m = 0; m < MAX_ACC; ++m)
        if (P->acc[m]) (void)hipFree(P->acc[m]);
    delete P;
    return PHX_OK;
}

int phx_plan_create(const phx_plan_desc* D, phx_plan** out) {
    if (!D || !out) return fail(PHX_E_INVALID, "null argument");
    *out = nullptr;
    if (D->abi_version != PHX_ABI_VERSION) return fail(PHX_E_INVALID, "ABI version mismatch");
    if (D->nblocks < 1 || !D->blocks) return fail(PHX_E_INVALID, "no blocks");
    if (D->ncases < 1) return fail(PHX_E_INVALID, "ncases must be >= 1");
    if (D->flux != PHX_FLUX_ROE && D->flux != PHX_FLUX_HLLL)
        return fail(PHX_E_INVALID, "Flux function type is not available.");
    if (D->limiter < PHX_LIMITER_VENKATAKRISHNAN || D->limiter > PHX_LIMITER_VAN_LEER)
        return fail(PHX_E_INVALID, "MUSCL: Slope limiter type not specified.");
    if (D->reconstruction != PHX_RECON_CONSERVATIVE && D->reconstruction != PHX_RECON_PRIMITIVE)
        return fail(PHX_E_INVALID, "unknown reconstruction type");
    if (D->quadrature_points < 1 || D->quadrature_points > PHX_MAX_QUADRATURE_POINTS)
        return fail(PHX_E_INVALID, "quadrature_points must be 1, 2 or 3");
    const int nq = D->quadrature_points;
    if (D->nstages < 1 || D->nstages > PHX_MAX_STAGES || !D->butcher_a)
        return fail(PHX_E_INVALID, "invalid Butcher table");
    if (!D->t_final) return fail(PHX_E_INVALID, "t_final is null");
    int64_t total = 0;
    for (int b = 0; b < D->nblocks; ++b) {
        const phx_block_desc& d = D->blocks[b];
        const std::string where = "block " + std::to_string(b) + ": ";
        if (d.nx < 1 || d.ny < 1) return fail(PHX_E_INVALID, where + "nx, ny must be >= 1");
        if (d.case_id < 0 || d.case_id >= D->ncases) return fail(PHX_E_INVALID, where + "case_id out of range");
        if (d.state_offset < 0 || d.state_offset % 4) return fail(PHX_E_INVALID, where + "state_offset must be a non-negative multiple of 4");
        if (!d.geom) return fail(PHX_E_INVALID, where + "geometry is null");
        int64_t glen;
        if (d.geom_kind == PHX_GEOM_CARTESIAN) glen = 8 + 11LL * d.nx + 11LL * d.ny + (int64_t)d.nx * d.ny;
        else if (d.geom_kind == PHX_GEOM_GENERAL) glen = 7LL * d.ny * (d.nx + 1) + 7LL * (d.ny + 1) * d.nx + 5LL * d.ny * d.nx;
        else return fail(PHX_E_INVALID, where + "unknown geometry kind");
        if (nq > 1) glen += 8LL * nq * d.nx * d.ny;  // Gauss-point offsets
        if (d.geom_len != glen) return fail(PHX_E_INVALID, where + "geometry length mismatch");
        for (int s = 0; s < 4; ++s) {
            const int st = d.side_type[s];
            if (st < PHX_SIDE_NEIGHBOR || st > PHX_SIDE_OUTLET) return fail(PHX_E_INVALID, where + "invalid side type");
            if (st == PHX_SIDE_NEIGHBOR) {
                const int n = d.neighbor[s];
                if (n < 0 || n >= D->nblocks) return fail(PHX_E_INVALID, where + "neighbour index out of range");
                const phx_block_desc& o = D->blocks[n];
                if ((s == PHX_E || s == PHX_W) && o.ny != d.ny) return fail(PHX_E_INVALID, where + "E/W neighbour ny mismatch");
                if ((s == PHX_N || s == PHX_S) && o.nx != d.nx) return fail(PHX_E_INVALID, where + "N/S neighbour nx mismatch");
                if (o.case_id != d.case_id) return fail(PHX_E_INVALID, where + "neighbour belongs to another case");
            }
        }
        total = std::max<int64_t>(total, d.state_offset + 4LL * d.nx * d.ny);
    }
    // overlap check
    {
        std::vector<std::pair<int64_t, int64_t>> iv;
        for (int b = 0; b < D->nblocks; ++b)
            iv.push_back({D->blocks[b].state_offset, D->blocks[b].state_offset + 4LL * D->blocks[b].nx * D->blocks[b].ny});
        std::sort(iv.begin(), iv.end());
        for (size_t i = 1; i < iv.size(); ++i)
            if (iv[i].first < iv[i - 1].second) return fail(PHX_E_INVALID, "block state ranges overlap");
    }

    phx_plan* P = new phx_plan();
    auto bail = [&](int rc) { phx_plan_destroy(P); return rc; };
    int rc;
#define PHX_CKP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) return bail(fail(PHX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_))); \
    } while (0)
    PHX_CKP(hipGetDevice(&P->device));
    PHX_CKP(hipDeviceGetAttribute(&P->ncu, hipDeviceAttributeMultiprocessorCount, P->device));
    P->nblocks = D->nblocks;
    P->ncases = D->ncases;
    P->nslot = dt_slots_for(D->ncases);
    P->flux = D->flux;
    P->ph = Phys{D->gamma, D->gm1, D->one_over_gm1, D->g_over_gm1, D->cfl, D->limiter,
                 D->reconstruction == PHX_RECON_PRIMITIVE ? 1 : 0, nq};
    P->desc.assign(D->blocks, D->blocks + D->nblocks);
    P->state_len = total;
    P->rows_per_item = D->rows_per_item > 0 ? D->rows_per_item : 32;
    P->grid_waves = D->grid_waves;

    // geometry upload (each block aligned to 4 doubles)
    std::vector<int64_t> goff(D->nblocks);
    int64_t gtot = 0;
    for (int b = 0; b < D->nblocks; ++b) {
        goff[b] = gtot;
        gtot += (D->blocks[b].geom_len + 3) / 4 * 4;
    }
    PHX_CKP(hipMalloc(&P->dgeom, std::max<int64_t>(gtot, 4) * sizeof(double)));
    for (int b = 0; b < D->nblocks; ++b)
        PHX_CKP(hipMemcpy(P->dgeom + goff[b], D->blocks[b].geom, D->blocks[b].geom_len * sizeof(double), hipMemcpyHostToDevice));
    P->hblocks.resize(D->nblocks);
    std::vector<long long> ghost_off(D->nblocks);
    long long gacc = 0;
    for (int b = 0; b < D->nblocks; ++b) {
        const phx_block_desc& d = D->blocks[b];
        DevBlock& h = P->hblocks[b];
        memset(&h, 0, sizeof(h));
        h.nx = d.nx; h.ny = d.ny; h.case_id = d.case_id; h.kind = d.geom_kind;
        for (int s = 0; s < 4; ++s) { h.side[s] = d.side_type[s]; h.nbr[s] = d.neighbor[s]; }
        h.off = d.state_offset;
        h.yoff = P->yf_len;
        P->yf_len += 4LL * (d.ny + 1) * d.nx;
        h.foff = P->nfaces;
        P->nfaces += (long long)d.ny * (d.nx + 1) + (long long)(d.ny + 1) * d.nx;
        h.geom = P->dgeom + goff[b];
        h.qoff = nq > 1 ? h.geom + (d.geom_len - 8LL * nq * d.nx * d.ny) : nullptr;
        if (d.geom_kind == PHX_GEOM_CARTESIAN)
            for (int q = 0; q < 8; ++q) h.wall[q] = d.geom[q];
        ghost_off[b] = gacc;
        gacc += 4LL * (2 * d.ny + 2 * d.nx);
    }
    PHX_CKP(hipMalloc(&P->dblocks, D->nblocks * sizeof(DevBlock)));
    PHX_CKP(hipMemcpy(P->dblocks, P->hblocks.data(), D->nblocks * sizeof(DevBlock), hipMemcpyHostToDevice));
    PHX_CKP(hipMalloc(&P->dghost_off, D->nblocks * sizeof(long long)));
    PHX_CKP(hipMemcpy(P->dghost_off, ghost_off.data(), D->nblocks * sizeof(long long), hipMemcpyHostToDevice));

    P->nstages = D->nstages;
    P->butcher.assign(D->butcher_a, D->butcher_a + (size_t)D->nstages * D->nstages);
    rc = build_programs(P, D->nstages, D->butcher_a);
    if (rc) return bail(rc);
    rc = build_items(P);
    if (rc) return bail(rc);
    if (nq > 1) {  // Gauss faces run a split stage: one y-face per wave, 3 cell rows per workgroup
        P->rows_b = 3;
        rc = phx_set_variant(P, 6, 0);
        if (rc) return bail(rc);
    }

    const size_t sbytes = (size_t)std::max<int64_t>(total, 4) * sizeof(double);
    PHX_CKP(hipMalloc(&P->sbuf[0], sbytes));
    if (D->nstages > 2) PHX_CKP(hipMalloc(&P->sbuf[1], sbytes));
    for (int m = 0; m < P->nacc; ++m) PHX_CKP(hipMalloc(&P->acc[m], sbytes));
    const size_t cb = (size_t)D->ncases * sizeof(double);
    PHX_CKP(hipMalloc(&P->dt, cb));
    PHX_CKP(hipMalloc(&P->t, cb));
    PHX_CKP(hipMalloc(&P->tfinal, cb));
    PHX_CKP(hipMalloc(&P->dtmin, (size_t)D->ncases * P->nslot * DT_STRIDE * sizeof(unsigned long long)));
    PHX_CKP(hipMalloc(&P->step, sizeof(int)));
    PHX_CKP(hipMalloc(&P->bad, D->ncases * sizeof(int)));
    PHX_CKP(hipMemcpy(P->tfinal, D->t_final, cb, hipMemcpyHostToDevice));
    PHX_CKP(hipMemset(P->dt, 0, cb));
    std::vector<double> zeros(D->ncases, 0.0);
    double* t0 = nullptr;
    PHX_CKP(hipMalloc(&t0, cb));
    PHX_CKP(hipMemcpy(t0, zeros.data(), cb, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_reset, dim3((D->ncases * P->nslot + 255) / 256), dim3(256), 0, 0, D->ncases, P->nslot, t0, P->t,
                       P->dtmin,
                       P->bad, P->step, 0);
    PHX_CKP(hipGetLastError());
    PHX_CKP(hipDeviceSynchronize());
    PHX_CKP(hipFree(t0));
    rc = phx::host::adjoint_prepare(P);
    if (rc) return bail(rc);
    PHX_CKP(hipStreamCreateWithFlags(&P->cap_stream, hipStreamNonBlocking));
    PHX_CKP(hipEventCreateWithFlags(&P->ev_a, hipEventDisableTiming));
    PHX_CKP(hipEventCreateWithFlags(&P->ev_b, hipEventDisableTiming));
    PHX_CKP(hipStreamCreateWithFlags(&P->side_stream, hipStreamNonBlocking));
    PHX_CKP(hipEventCreateWithFlags(&P->ev_fork, hipEventDisableTiming));
    PHX_CKP(hipEventCreateWithFlags(&P->ev_join, hipEventDisableTiming));
#undef PHX_CKP
    *out = P;
    return PHX_OK;
}

int phx_reset(phx_plan* P, const double* t0_host, int32_t step0, void* stream) {
    if (!P || !t0_host) return fail(PHX_E_INVALID, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const size_t cb = (size_t)P->ncases * sizeof(double);
    double* t0 = nullptr;
    PHX_CK(hipMalloc(&t0, cb));
    PHX_CK(hipMemcpy(t0, t0_host, cb, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_reset, dim3((P->ncases * P->nslot + 255) / 256), dim3(256), 0, st, P->ncases, P->nslot, t0,
                       P->t, P->dtmin,
                       P->bad, P->step, step0);
    PHX_CK(hipGetLastError());
    PHX_CK(hipStreamSynchronize(st));
    PHX_CK(hipFree(t0));
    return PHX_OK;
}

int phx_run(phx_plan* P, double* U, int32_t nsteps, int32_t use_graph, void* stream) {
    if (!P || !U) return fail(PHX_E_INVALID, "null argument");
    if (nsteps < 0) return fail(PHX_E_INVALID, "nsteps must be >= 0");
    hipStream_t st = (hipStream_t)stream;
    launch_dt_reduce(P, U, st);
    PHX_CK(hipGetLastError());
    if (!use_graph) {
        for (int s = 0; s < nsteps; ++s) launch_step(P, U, st, nullptr);
        PHX_CK(hipGetLastError());
        return PHX_OK;
    }
    if (P->graph_U != U) release_graphs(P);
    // GRAPH_STEPS steps per replay while enough remain (one graph launch per GRAPH_STEPS steps),
    // then single steps
    const int ng = nsteps / GRAPH_STEPS, n1 = nsteps % GRAPH_STEPS;
    for (int g = 0; g < 2; ++g) {  // both captured on first use, so that later calls only replay
        if (P->gexec[g]) continue;
        PHX_CK(hipStreamBeginCapture(P->cap_stream, hipStreamCaptureModeThreadLocal));
        for (int s = 0; s < (g == 0 ? 1 : GRAPH_STEPS); ++s) launch_step(P, U, P->cap_stream, nullptr);
        PHX_CK(hipStreamEndCapture(P->cap_stream, &P->graph[g]));
        PHX_CK(hipGraphInstantiate(&P->gexec[g], P->graph[g], nullptr, nullptr, 0));
    }
    P->graph_U = U;
    PHX_CK(hipEventRecord(P->ev_a, st));
    PHX_CK(hipStreamWaitEvent(P->cap_stream, P->ev_a, 0));
    for (int s = 0; s < ng; ++s) PHX_CK(hipGraphLaunch(P->gexec[1], P->cap_stream));
    for (int s = 0; s < n1; ++s) PHX_CK(hipGraphLaunch(P->gexec[0], P->cap_stream));
    PHX_CK(hipEventRecord(P->ev_b, P->cap_stream));
    PHX_CK(hipStreamWaitEvent(st, P->ev_b, 0));
    return PHX_OK;
}

int phx_run_frozen(phx_plan* P, double* U, int32_t nsteps, const double* dt_sched, void* stream) {
    if (!P || !U || (!dt_sched && nsteps > 0)) return fail(PHX_E_INVALID, "null argument");
    hipStream_t st = (hipStream_t)stream;
    for (int s = 0; s < nsteps; ++s) launch_step(P, U, st, dt_sched + (size_t)s * P->ncases);
    PHX_CK(hipGetLastError());
    return PHX_OK;
}

int phx_run_record(phx_plan* P, double* U, int32_t nsteps, const double* dt_sched, double* stages, void* stream) {
    if (!P || !U || ((!dt_sched || !stages) && nsteps > 0)) return fail(PHX_E_INVALID, "null argument");
    if (nsteps < 0) return fail(PHX_E_INVALID, "nsteps must be >= 0");
    hipStream_t st = (hipStream_t)stream;
    const size_t len = (size_t)P->state_len;
    const int S = (int)P->progs.size();
    if (nsteps == 0) return PHX_OK;
    // The records are the working buffers: step n reads its input from record (n, 0), writes the
    // state of stage s+1 into record (n, s+1) and its result into record (n+1, 0) (U after the last
    // step), so only the first input is copied.
    PHX_CK(hipMemcpyAsync(stages, U, len * sizeof(double), hipMemcpyDeviceToDevice, st));
    for (int n = 0; n < nsteps; ++n) {
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN_THREADS), 0, st, P->ncases, P->nslot, P->ph.cfl, P->dtmin, P->tfinal, P->t,
                           P->dt, P->step, P->bad, dt_sched + (size_t)n * P->ncases, 1);
        double* rec = stages + (size_t)n * S * len;
        double* next = (n + 1 < nsteps) ? rec + (size_t)S * len : U;
        for (int s = 0; s < S; ++s) {
            BufMap M;
            M.u_read = rec;
            M.u_write = next;
            M.s_read = rec + (size_t)s * len;
            M.s_write = (s + 1 < S) ? rec + (size_t)(s + 1) * len : next;  // one-stage tables write S0
            launch_prog_stage_mapped(P, (size_t)s, M, st);
        }
    }
    PHX_CK(hipGetLastError());
    return PHX_OK;
}

int phx_stage(phx_plan* P, double* U, int32_t stage, void* stream) {
    if (!P || !U) return fail(PHX_E_INVALID, "null argument");
    if (stage < 0 || stage >= (int)P->progs.size()) return fail(PHX_E_INVALID, "stage index out of range");
    launch_prog_stage(P, U, (size_t)stage, (hipStream_t)stream);
    PHX_CK(hipGetLastError());
    return PHX_OK;
}

int phx_residual(phx_plan* P, const double* U, double* R, void* stream) {
    if (!P || !U || !R) return fail(PHX_E_INVALID, "null argument");
    StageArgs A = base_args(P);
    A.Uin = U;
    A.Rout = R;
    A.nout = 0;
    launch_stage(P, A, (hipStream_t)stream);
    PHX_CK(hipGetLastError());
    return PHX_OK;
}

int phx_dt(phx_plan* P, const double* U, double* dt_dev, void* stream) {
    if (!P || !U || !dt_dev) return fail(PHX_E_INVALID, "null argument");
    hipStream_t st = (hipStream_t)stream;
    launch_dt_reduce(P, U, st);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN_THREADS), 0, st, P->ncases, P->nslot, P->ph.cfl, P->dtmin, P->tfinal, P->t,
                       dt_dev, P->step, P->bad, (const double*)nullptr, 0);
    PHX_CK(hipGetLastError());
    return PHX_OK;
}

int phx_ghosts(phx_plan* P, const double* U, double* G, void* stream) {
    if (!P || !U || !G) return fail(PHX_E_INVALID, "null argument");
    hipLaunchKernelGGL(k_ghosts, dim3(P->nblocks), dim3(256), 0, (hipStream_t)stream, P->dblocks, P->nblocks,
                       P->dghost_off, U, G);
    PHX_CK(hipGetLastError());
    return PHX_OK;
}

int phx_status(phx_plan* P, int32_t* bad_step, int32_t* step, double* t_host, double* dt_host) {
    if (!P) return fail(PHX_E_INVALID, "null argument");
    PHX_CK(hipDeviceSynchronize());
    std::vector<int> b(P->ncases);
    int s = 0;
    PHX_CK(hipMemcpy(b.data(), P->bad, P->ncases * sizeof(int), hipMemcpyDeviceToHost));
    PHX_CK(hipMemcpy(&s, P->step, sizeof(int), hipMemcpyDeviceToHost));
    if (bad_step)
        for (int c = 0; c < P->ncases; ++c) bad_step[c] = (b[c] == 0x7fffffff) ? -1 : b[c];
    if (step) *step = s;
    if (t_host) PHX_CK(hipMemcpy(t_host, P->t, P->ncases * sizeof(double), hipMemcpyDeviceToHost));
    if (dt_host) PHX_CK(hipMemcpy(dt_host, P->dt, P->ncases * sizeof(double), hipMemcpyDeviceToHost));
    return PHX_OK;
}

int phx_set_launch(phx_plan* P, int32_t rows_per_item, int32_t grid_waves, int32_t waves_per_simd) {
    if (!P || rows_per_item < 1 || grid_waves < 0 || waves_per_simd < 0 || waves_per_simd > 4 || waves_per_simd == 1)
        return fail(PHX_E_INVALID, "invalid launch parameters");
    PHX_CK(hipDeviceSynchronize());
    P->rows_per_item = rows_per_item;
    P->grid_waves = grid_waves;
    P->waves_per_simd = waves_per_simd == 0 ? 2 : waves_per_simd;
    return build_items(P);
}

int phx_set_variant(phx_plan* P, int32_t variant, int32_t rows_b) {
    if (!P || variant < 0 || variant > 8 || rows_b < 0) return fail(PHX_E_INVALID, "invalid stage variant");
    if (P->ph.nq > 1 && variant != 1 && variant != 5 && variant != 6)
        return fail(PHX_E_INVALID, "more than one Gauss point per face runs the split stages only (variants 1, 5, 6)");
    if (variant == 8) {
        const int rb = rows_b > 0 ? rows_b : P->rows_b;
        const int T = (std::min(std::max(rb, 1), FLAT_MAX_WAVES - 1) + 1) * WAVE;
        for (int b = 0; b < P->nblocks; ++b)
            if ((P->desc[b].nx + 64 + T - 1) / T * T > FLAT_MAX_FACES)
                return fail(PHX_E_INVALID, "a block is too wide for the flat y-face stage (variant 8: nx + 64 faces per range)");
    }
    PHX_CK(hipDeviceSynchronize());
    if ((variant == 1 || variant >= 5) && !P->sn) {
        const size_t L = (size_t)std::max<int64_t>(P->state_len, 4);
        PHX_CK(hipMalloc(&P->sn, 2 * (size_t)std::max(P->ph.nq, 1) * L * sizeof(double)));
        PHX_CK(hipMalloc(&P->p1, L * sizeof(double)));
        PHX_CK(hipMalloc(&P->yf, (size_t)std::max<long long>(P->yf_len, 4) * sizeof(double)));
    }
    if (variant == 3 && !P->ew_rb) {
        const size_t cells = (size_t)std::max<int64_t>(P->state_len, 4) / 4;
        PHX_CK(hipMalloc(&P->ew_rb, phx::adj::RB * cells * sizeof(double)));
        PHX_CK(hipMalloc(&P->ew_ff, 4 * (size_t)std::max<long long>(P->nfaces, 1) * sizeof(double)));
    }
    P->variant = variant;
    if (rows_b > 0) P->rows_b = rows_b;
    return build_items(P);
}

int phx_set_xcd_order(phx_plan* P, int32_t chunk) {
    if (!P || chunk < 0) return fail(PHX_E_INVALID, "invalid XCD chunk (0 or a positive number of work items)");
    PHX_CK(hipDeviceSynchronize());
    P->xcd_chunk = chunk;
    return build_items(P);
}

int phx_launch_info(const phx_plan* P, int32_t* items_cart, int32_t* items_gen, int32_t* grid_cart, int32_t* grid_gen) {
    if (!P) return fail(PHX_E_INVALID, "null argument");
    if (items_cart) *items_cart = (int32_t)P->items[1].size();
    if (items_gen) *items_gen = (int32_t)P->items[0].size();
    if (grid_cart) *grid_cart = P->grid[1];
    if (grid_gen) *grid_gen = P->grid[0];
    return PHX_OK;
}

}  // extern "C"
