// phx_plan.hpp -- host-side plan object shared by the stage and adjoint translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/pyhype_hip.h"
#include "phx_device.hpp"

namespace phx {
namespace host {

constexpr int MAX_OUT = 8;  // outputs of one stage program (stage state + accumulators)
constexpr int MAX_ACC = 8;  // accumulator buffers of the RK programs

// One work item of the stage kernel: block, first owned column, row range [r0, r1).  Tiles of the
// element-wise stage use the same record as {block, first element, element count, 0}.
struct WorkItem {
    int block, j0, r0, r1;
};

enum BufId { BUF_USER = 0, BUF_S0 = 1, BUF_S1 = 2, BUF_ACC0 = 3 };

struct StageOut {
    int dst, src;
    double a;
};
struct StageProg {
    int in;
    std::vector<StageOut> outs;
    int state_out;
};

// Records `msg` as the thread's last error (phx_last_error) and returns `code`.
int fail(int code, const std::string& msg);

// Per-block offsets of the adjoint work arrays: cells (in cells) and faces (in faces; the
// ny*(nx+1) x-faces of the block first, then its (ny+1)*nx y-faces).
struct AdjBlock {
    long long cell, face;
};

// Launch tile of the element-wise adjoint kernels: `count` consecutive elements of one block.
struct AdjTile {
    int block, start, count;
};

}  // namespace host
}  // namespace phx

#define PHX_CK(call)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return ::phx::host::fail(PHX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct phx_plan {
    int device = 0;
    int ncu = 0;
    // split and element-wise stages: work items dealt to the XCDs in chunks of xcd_chunk (0: one
    // contiguous range per XCD)
    int xcd_chunk = 0;
    int nblocks = 0, ncases = 0, flux = 0;
    phx::Phys ph{};
    std::vector<phx_block_desc> desc;
    std::vector<phx::dev::DevBlock> hblocks;
    phx::dev::DevBlock* dblocks = nullptr;
    double* dgeom = nullptr;
    int64_t state_len = 0;
    int rows_per_item = 32, grid_waves = 0, waves_per_simd = 2;
    std::vector<phx::host::WorkItem> items[2];  // [0] general, [1] cartesian
    phx::host::WorkItem* ditems[2] = {nullptr, nullptr};
    int grid[2] = {0, 0};
    // stage variant: 0 = fused rolling rows, 1 = split (reconstruct + x-faces, y-faces, update),
    // 2 = shared-halo workgroups, 3 = element-wise (cell-variable lanes, face lanes, cell lanes),
    // 4 = paired waves, 5 = split with y-faces and update fused, 6 = split with one y-face per wave,
    // 7 = variant 6 with the row-major reconstruction kernel, 8 = the row-major reconstruction and
    // y-faces and update over flat cell ranges
    int variant = 0, rows_b = 8;
    std::vector<phx::host::WorkItem> items_b[2];  // split variant: update kernel
    int flat_q[2] = {0, 0};                          // variant 8: LDS faces per lane of the update kernel
    phx::host::WorkItem* ditems_b[2] = {nullptr, nullptr};
    int grid_b[2] = {0, 0};
    std::vector<phx::host::WorkItem> items_f[2];  // variant 1: y-face kernel; variant 7: reconstruction
    phx::host::WorkItem* ditems_f[2] = {nullptr, nullptr};
    int grid_f[2] = {0, 0};
    // element-wise variant: tiles of [0] the reconstruction (cells), [1] the faces, [2] the update (cells)
    std::vector<phx::host::WorkItem> items_e[3][2];
    phx::host::WorkItem* ditems_e[3][2] = {};
    // element-wise stage on a mesh with both geometry kinds: each phase's items of both kinds in one
    // list (one launch; the kernels pick the geometry path per block)
    phx::host::WorkItem* ditems_em[3] = {};
    int n_em[3] = {0, 0, 0}, grid_em[3] = {0, 0, 0};
    // split stages 6 and 7 likewise: merged lists of items (variant 6 reconstruction), items_b (y-faces
    // and update) and items_f (variant 7 reconstruction)
    phx::host::WorkItem* ditems_m[3] = {};
    int n_m[3] = {0, 0, 0}, grid_m[3] = {0, 0, 0};
    int grid_e[3][2] = {};
    double* ew_rb = nullptr;  // element-wise variant: limited face states E, W, N, S per cell (16 doubles)
    double* ew_ff = nullptr;  // element-wise variant: face fluxes (4 doubles per face, DevBlock::foff)
    long long nfaces = 0;     // faces of all blocks (x-faces and y-faces)
    double* sn = nullptr;  // split variant: N/S states per cell
    double* p1 = nullptr;  // split variant: x-face part of the residual per cell
    double* yf = nullptr;  // split variant: y-face fluxes
    long long yf_len = 0;  // doubles of yf (sum over blocks of (ny+1)*nx*4)
    std::vector<phx::host::StageProg> progs;
    std::vector<double> butcher;  // nstages x nstages, row-major
    int nstages = 0;
    int nacc = 0;
    double* sbuf[2] = {nullptr, nullptr};
    double* acc[phx::host::MAX_ACC] = {};
    double *dt = nullptr, *t = nullptr, *tfinal = nullptr;
    unsigned long long* dtmin = nullptr;  // nslot partial CFL minima per case, one cache line each
    int nslot = 1;
    int *step = nullptr, *bad = nullptr;
    long long* dghost_off = nullptr;
    hipStream_t cap_stream = nullptr;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    // the second geometry kind of a phase runs on side_stream, concurrently with the first
    // (fork/join through ev_fork/ev_join; captured into the step graphs as parallel branches)
    hipStream_t side_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // captured steps: [0] one step, [1] GRAPH_STEPS steps (replayed by phx_run)
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    double* graph_U = nullptr;
    bool copy_back = false;  // one-stage tables: result lands in sbuf[0] and is copied to U

    // discrete adjoint: per-block offsets and launch tiles, built when the plan is created (when its
    // neighbour links are mutual; otherwise adj_meta_error says why not)
    bool adj_meta_ready = false;
    std::string adj_meta_error;
    // discrete adjoint work arrays, allocated on the first adjoint call
    bool adj_ready = false;
    long long adj_cells = 0, adj_faces = 0;
    phx::host::AdjBlock* dadj = nullptr;
    std::vector<phx::host::AdjTile> adj_cell_tiles, adj_face_tiles;
    phx::host::AdjTile *dadj_cell_tiles = nullptr, *dadj_face_tiles = nullptr;
    double* adj_rb = nullptr;     // limited face states E, W, N, S of every cell (16 doubles)
    double* adj_face = nullptr;   // adjoints of the left/right states of every face (8 doubles)
    double* adj_slot = nullptr;   // reconstruction adjoints: cell, W, E, S, N (20 doubles per cell)
    double* adj_lam[PHX_MAX_STAGES] = {};  // stage-state adjoints of one step (stages 1..S-1)
    // tangent-linear products, allocated on the first tangent call
    double* tan_k[PHX_MAX_STAGES] = {};    // stage residual derivatives of one step
    double* tan_u = nullptr;               // derivative of the current stage input
};

namespace phx {
namespace host {
// Frees the adjoint work arrays of a plan (no-op when none were allocated).
void adjoint_release(phx_plan* P);
// Set-up of a new plan's adjoint: offsets and launch tiles (when the neighbour links are mutual),
// and the address of the static work array of phx_residual_vjp_noalloc on the plan's device.
int adjoint_prepare(phx_plan* P);
}  // namespace host
}  // namespace phx
