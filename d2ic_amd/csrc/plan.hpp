// Device layout of a solver plan (d2ic_cg_plan), shared by the solver and the
// adjoint kernels.
//
//   * pixels grouped by element and cut into tiles of <= kTile pixels; pixel i
//     of tile t sits at slot t * kTile + i (padded slots: centre (0.5, 0.5),
//     zero shape functions); per slot px, py, N4 (the 4 shape values) and F
//     (f at the pixel);
//   * per tile: tile_cnt, tile_nodes (its element's 4 node ids), TV (static
//     record, kTV ints: [0] pixel count, [1..4] floor(min x) - 1, floor(max x),
//     floor(min y) - 1, floor(max y), [5] 1 if every pixel sits at an exact
//     pixel centre (j + 0.5, i + 0.5)), TR (kTR words: col0, row0, log2 of
//     the row width, ax, bx, ay, by, regular flag -- a regular tile's pixels
//     fill rows of 2^k columns from (col0, row0) and their shape values are
//     exactly the Q4 functions of xi = ax col + bx, eta = ay row + by, so the
//     pixel pass loads only F for it);
//   * node -> (tile, local node) gather table: node_ptr (nn + 1), node_ent
//     entries (t << 2) | a;
//   * spring stencil: nb_idx (nn x nb_maxdeg, -1 padded) and nb_wp = the
//     combined weight w_ij (c_i + c_j) / 2.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace d2ic {
constexpr int kTile = 256;  // pixels per tile (one element, or a 256-pixel run of one)
constexpr int kTV = 8;
constexpr int kTR = 8;  // regular-geometry record per tile
}  // namespace d2ic

struct d2ic_cg_plan {
    int device = 0;
    int64_t np = 0, nn = 0, ntiles = 0, nchunks = 0;
    int nb_maxdeg = 0;
    int32_t* tile_cnt = nullptr;
    int4* tile_nodes = nullptr;
    float* px = nullptr;
    float* py = nullptr;
    float4* N4 = nullptr;
    float* F = nullptr;
    int32_t* node_ptr = nullptr;
    int32_t* node_ent = nullptr;
    int32_t* nb_idx = nullptr;
    float* nb_wp = nullptr;
    float* U1 = nullptr;
    float* G[2] = {nullptr, nullptr};
    float* D[2] = {nullptr, nullptr};
    float* TP = nullptr;
    float* TV = nullptr;
    float* TR = nullptr;  // per tile: col0, row0, log2 rw, ax, bx, ay, by, regular flag
    double* NCP = nullptr;
    int2* item_tr = nullptr;  // per node item: first and last tile touching its nodes
    int ref_H = 0, ref_W = 0, interp = 0;
    bool has_ref = false;
    size_t bytes = 0;
    std::vector<void*> allocs;
    // start/end of the last persistent-kernel launch this plan led (plans[0])
    hipEvent_t ev_start = nullptr, ev_end = nullptr;
    bool timed = false;
    // persistent grid size for single-problem solves led by this plan (0: the
    // work-based default), set by d2ic_cg_plan_set_grid (the autotuner)
    int grid_blocks = 0;
    // tiling chosen at creation (d2ic_cg_plan_create_tiled): pixels per tile
    // at most, and the column-strip width of elements wider than 24 px
    int tile_px = 256, strip_w = 16;
};

namespace d2ic {
// Trivially copyable view of a plan's device layout (kernel argument).
struct PlanView {
    int64_t np, nn, ntiles;
    int nb_maxdeg, interp;
    const int32_t* tile_cnt;
    const int4* tile_nodes;
    const float* px;
    const float* py;
    const float4* N4;
    const float* F;
    const int32_t* node_ptr;
    const int32_t* node_ent;
    const int32_t* nb_idx;
    const float* nb_wp;
};

inline PlanView plan_view(const d2ic_cg_plan* p) {
    PlanView v;
    v.np = p->np;
    v.nn = p->nn;
    v.ntiles = p->ntiles;
    v.nb_maxdeg = p->nb_maxdeg;
    v.interp = p->interp;
    v.tile_cnt = p->tile_cnt;
    v.tile_nodes = p->tile_nodes;
    v.px = p->px;
    v.py = p->py;
    v.N4 = p->N4;
    v.F = p->F;
    v.node_ptr = p->node_ptr;
    v.node_ent = p->node_ent;
    v.nb_idx = p->nb_idx;
    v.nb_wp = p->nb_wp;
    return v;
}
}  // namespace d2ic
