// Local Gauss-Newton nodal refinement (reference: d2ic/solver_local_gn.py).
//
// n_sweeps relaxed Jacobi sweeps, each:
//   * pixel pass (_pixel_state, :205-242): u at the pixel from its 4 nodes,
//     r = g(x + u) - f(x), and the central-difference gradient images of g
//     sampled at x + u (the gradient images, not the interpolant's
//     derivative: :186-202);
//   * node pass (jacobi_nodal_step_spring, :312-414): per node, over its
//     pixels in the node->pixel table, J = (gx, gy) N, g = sum r J,
//     H = sum J J^T + lam tr(H) I, plus the spring term alpha w_reg
//     sum_j w (u_n - u_j) with alpha w_reg sum_j w I; delta = -H^-1 g (with
//     1e-8 I), clipped to |delta| <= max_step, u += omega delta.
// History per sweep: RMS residual over the pixels and RMS displacement update.
//
// Two layouts:
//   * d2ic_lgn_solve: the reference's own pixel order and node->pixel table
//     (as built by build_pixel_assets); one thread per pixel, one wave per
//     node walking the node's table row (20 B per entry, 4 entries per pixel);
//   * d2ic_lgn_solve_plan: the global solver's plan (pixels grouped by element
//     into tiles of <= 256); one block per tile reduces the tile's pixels to
//     5 sums (g0, g1, h00, h01, h11) per tile node, and a node gathers its
//     <= 4 (or 16 for split elements) tile sums through the plan's node->tile
//     table, so the node pass reads ~20 B per tile entry instead of per pixel.
// Sampling positions use the split integer-cell + fraction form of the global
// solver (more accurate than x + u in FP32 at large x).
#include <algorithm>

#include "common.hpp"
#include "interp.hpp"
#include "plan.hpp"

namespace d2ic {

namespace {

constexpr int kLgnBlock = 256;
constexpr int kLgnNodesPerBlock = kLgnBlock / kWave;

// gx[:, 1:-1] = 0.5 (im[:, 2:] - im[:, :-2]); gy likewise along rows; edges 0.
__global__ void k_lgn_grad(const float* __restrict__ img, int H, int W, float* __restrict__ gx, float* __restrict__ gy) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)H * W) return;
    const int y = (int)(i / W), x = (int)(i % W);
    gx[i] = (x >= 1 && x < W - 1) ? 0.5f * (img[i + 1] - img[i - 1]) : 0.f;
    gy[i] = (y >= 1 && y < H - 1) ? 0.5f * (img[i + W] - img[i - W]) : 0.f;
}

__device__ __forceinline__ float sample_value(int interp, const float* __restrict__ img, int H, int W, int x0, float fx,
                                              int y0, float fy) {
    float v, gx, gy;
    sample_at(interp, img, H, W, x0, fx, y0, fy, v, gx, gy);
    return v;
}

// f(x_p) once per solve.
__global__ void k_lgn_ref(const float2* __restrict__ xy, int64_t np, const float* __restrict__ img, int H, int W,
                          int interp, float* __restrict__ I1) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int x0, y0;
    float fx, fy;
    split_coord(xy[p].x, 0.f, x0, fx);
    split_coord(xy[p].y, 0.f, y0, fy);
    I1[p] = sample_value(interp, img, H, W, x0, fx, y0, fy);
}

// Block sum of a per-thread FP64 value in a fixed order; thread 0 stores it.
__device__ __forceinline__ void block_sum_store(double v, double* __restrict__ out) {
    __shared__ double s[kLgnBlock / kWave];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    v = wave_sum63(v);
    if (lane == kWave - 1) s[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kLgnBlock / kWave; ++w) t += s[w];
        *out = t;
    }
}

__global__ __launch_bounds__(kLgnBlock) void k_lgn_pixels(const float2* __restrict__ xy, const int4* __restrict__ pn,
                                                          const float4* __restrict__ N, int64_t np,
                                                          const float2* __restrict__ u, const float* __restrict__ img,
                                                          const float* __restrict__ gximg, const float* __restrict__ gyimg,
                                                          int H, int W, int interp, const float* __restrict__ I1,
                                                          float* __restrict__ r, float* __restrict__ gxd,
                                                          float* __restrict__ gyd, double* __restrict__ r2_part) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double r2 = 0.0;
    if (p < np) {
        const int4 q = pn[p];
        const float4 w = N[p];
        const float2 ua = u[q.x], ub = u[q.y], uc = u[q.z], ud = u[q.w];
        const float ux = w.x * ua.x + w.y * ub.x + w.z * uc.x + w.w * ud.x;
        const float uy = w.x * ua.y + w.y * ub.y + w.z * uc.y + w.w * ud.y;
        int x0, y0;
        float fx, fy;
        split_coord(xy[p].x, ux, x0, fx);
        split_coord(xy[p].y, uy, y0, fy);
        const float rv = sample_value(interp, img, H, W, x0, fx, y0, fy) - I1[p];
        r[p] = rv;
        gxd[p] = sample_value(interp, gximg, H, W, x0, fx, y0, fy);
        gyd[p] = sample_value(interp, gyimg, H, W, x0, fx, y0, fy);
        r2 = (double)rv * (double)rv;
    }
    block_sum_store(r2, r2_part + blockIdx.x);
}

// The 2x2 update of node n from its image sums (g, H) and the spring term
// (jacobi_nodal_step_spring, solver_local_gn.py:312-414); returns |step|^2.
__device__ __forceinline__ double lgn_node_update(int64_t n, const float2* __restrict__ u, float g0, float g1, float h00,
                                                  float h01, float h11, bool has_pixels,
                                                  const int32_t* __restrict__ nbi, const int32_t* __restrict__ nbd,
                                                  const float* __restrict__ nbw, const float* __restrict__ nrw,
                                                  int64_t mdn, float lam, float max_step, float alpha, float omega,
                                                  float2* __restrict__ u_next) {
    const float2 ui = u[n];
    const int dn = nbd[n];
    float gr0 = 0.f, gr1 = 0.f, sumw = 0.f;
    for (int k = 0; k < dn; ++k) {
        const float w = nbw[n * mdn + k];
        const float2 uj = u[nbi[n * mdn + k]];
        gr0 += w * (ui.x - uj.x);
        gr1 += w * (ui.y - uj.y);
        sumw += w;
    }
    const float alpha_loc = alpha * nrw[n];
    const float lam_loc = lam * (h00 + h11 + 1e-12f);
    const float hreg = alpha_loc * sumw;
    const float a = h00 + lam_loc + hreg + 1e-8f, b = h01, d = h11 + lam_loc + hreg + 1e-8f;
    const float q0 = g0 + alpha_loc * gr0, q1 = g1 + alpha_loc * gr1;
    const bool has_update = has_pixels || (alpha != 0.f && dn > 0);
    float d0 = 0.f, d1 = 0.f;
    if (has_update) {
        const float det = a * d - b * b;
        d0 = -(d * q0 - b * q1) / det;
        d1 = -(a * q1 - b * q0) / det;
        const float nrm = sqrtf(d0 * d0 + d1 * d1);
        const float factor = fminf(1.0f, max_step / (nrm + 1e-12f));
        d0 *= factor;
        d1 *= factor;
    }
    const float2 un = make_float2(ui.x + omega * d0, ui.y + omega * d1);
    u_next[n] = un;
    const float sx = un.x - ui.x, sy = un.y - ui.y;
    return (double)sx * (double)sx + (double)sy * (double)sy;
}

// One wave per node: lanes stride over the node's pixels (FP32 sums, DPP
// reduction); the last lane adds the spring term and solves the 2x2 system.
__global__ __launch_bounds__(kLgnBlock) void k_lgn_nodes(
    const float2* __restrict__ u, const float* __restrict__ r, const float* __restrict__ gxd,
    const float* __restrict__ gyd, const int32_t* __restrict__ npi, const float* __restrict__ npw,
    const int32_t* __restrict__ npd, int64_t mdp, const int32_t* __restrict__ nbi, const int32_t* __restrict__ nbd,
    const float* __restrict__ nbw, const float* __restrict__ nrw, int64_t mdn, int64_t nn, float lam, float max_step,
    float alpha, float omega, float2* __restrict__ u_next, double* __restrict__ s2_part) {
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int64_t n = (int64_t)blockIdx.x * kLgnNodesPerBlock + wave;
    float g0 = 0.f, g1 = 0.f, h00 = 0.f, h01 = 0.f, h11 = 0.f;
    const int deg = n < nn ? npd[n] : 0;
    for (int k = lane; k < deg; k += kWave) {
        const int p = npi[n * mdp + k];
        const float Ni = npw[n * mdp + k];
        const float ri = r[p];
        const float jx = gxd[p] * Ni, jy = gyd[p] * Ni;
        g0 += ri * jx;
        g1 += ri * jy;
        h00 += jx * jx;
        h01 += jx * jy;
        h11 += jy * jy;
    }
    g0 = wave_sum63(g0);
    g1 = wave_sum63(g1);
    h00 = wave_sum63(h00);
    h01 = wave_sum63(h01);
    h11 = wave_sum63(h11);
    double s2 = 0.0;
    if (lane == kWave - 1 && n < nn)
        s2 = lgn_node_update(n, u, g0, g1, h00, h01, h11, deg > 0, nbi, nbd, nbw, nrw, mdn, lam, max_step, alpha, omega,
                             u_next);
    block_sum_store(s2, s2_part + blockIdx.x);
}

// ---- plan (tile) layout ----
static_assert(kTile == kLgnBlock, "one thread per tile slot");

// f at every tile slot (padded slots are never read)
__global__ void k_lgn_ref_tiles(PlanView pv, const float* __restrict__ img, int H, int W, int interp,
                                float* __restrict__ I1) {
    const int64_t t = blockIdx.x;
    const int i = threadIdx.x;
    if (i >= pv.tile_cnt[t]) return;
    const int64_t s = t * kTile + i;
    int x0, y0;
    float fx, fy;
    split_coord(pv.px[s], 0.f, x0, fx);
    split_coord(pv.py[s], 0.f, y0, fy);
    I1[s] = sample_value(interp, img, H, W, x0, fx, y0, fy);
}

// One block per tile: the pixel state (as k_lgn_pixels) and the tile's sums
// (g0, g1, h00, h01, h11) for each of its 4 nodes, reduced in a fixed order
// (wave DPP sums, then the 4 waves in order); r^2 per tile in FP64.
__global__ __launch_bounds__(kLgnBlock) void k_lgn_tiles(PlanView pv, const float2* __restrict__ u,
                                                         const float* __restrict__ img, const float* __restrict__ gximg,
                                                         const float* __restrict__ gyimg, int H, int W, int interp,
                                                         const float* __restrict__ I1, float* __restrict__ tsum,
                                                         double* __restrict__ r2_part) {
    __shared__ float s_v[kLgnBlock / kWave][20];
    __shared__ double s_r2[kLgnBlock / kWave];
    const int64_t t = blockIdx.x;
    const int i = threadIdx.x, lane = i & (kWave - 1), wave = i / kWave;
    const int4 q = pv.tile_nodes[t];
    float v[20];
#pragma unroll
    for (int k = 0; k < 20; ++k) v[k] = 0.f;
    double r2 = 0.0;
    if (i < pv.tile_cnt[t]) {
        const int64_t s = t * kTile + i;
        const float4 w = pv.N4[s];
        const float2 ua = u[q.x], ub = u[q.y], uc = u[q.z], ud = u[q.w];
        const float ux = w.x * ua.x + w.y * ub.x + w.z * uc.x + w.w * ud.x;
        const float uy = w.x * ua.y + w.y * ub.y + w.z * uc.y + w.w * ud.y;
        int x0, y0;
        float fx, fy;
        split_coord(pv.px[s], ux, x0, fx);
        split_coord(pv.py[s], uy, y0, fy);
        const float rv = sample_value(interp, img, H, W, x0, fx, y0, fy) - I1[s];
        const float gx = sample_value(interp, gximg, H, W, x0, fx, y0, fy);
        const float gy = sample_value(interp, gyimg, H, W, x0, fx, y0, fy);
        const float Na[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const float jx = gx * Na[a], jy = gy * Na[a];
            v[5 * a + 0] = rv * jx;
            v[5 * a + 1] = rv * jy;
            v[5 * a + 2] = jx * jx;
            v[5 * a + 3] = jx * jy;
            v[5 * a + 4] = jy * jy;
        }
        r2 = (double)rv * (double)rv;
    }
#pragma unroll
    for (int k = 0; k < 20; ++k) v[k] = wave_sum63(v[k]);
    r2 = wave_sum63(r2);
    if (lane == kWave - 1) {
#pragma unroll
        for (int k = 0; k < 20; ++k) s_v[wave][k] = v[k];
        s_r2[wave] = r2;
    }
    __syncthreads();
    if (i < 20) {
        float acc = 0.f;
        for (int w2 = 0; w2 < kLgnBlock / kWave; ++w2) acc += s_v[w2][i];
        tsum[t * 20 + i] = acc;
    } else if (i == 20) {
        double acc = 0.0;
        for (int w2 = 0; w2 < kLgnBlock / kWave; ++w2) acc += s_r2[w2];
        r2_part[t] = acc;
    }
}

// One thread per node: the node's tile sums gathered in the table's order,
// then the 2x2 update.
__global__ __launch_bounds__(kLgnBlock) void k_lgn_nodes_tiles(
    PlanView pv, const float2* __restrict__ u, const float* __restrict__ tsum, const int32_t* __restrict__ nbi,
    const int32_t* __restrict__ nbd, const float* __restrict__ nbw, const float* __restrict__ nrw, int64_t mdn,
    int64_t nn, float lam, float max_step, float alpha, float omega, float2* __restrict__ u_next,
    double* __restrict__ s2_part) {
    const int64_t n = (int64_t)blockIdx.x * kLgnBlock + threadIdx.x;
    double s2 = 0.0;
    if (n < nn) {
        const int e0 = pv.node_ptr[n], e1 = pv.node_ptr[n + 1];
        float g0 = 0.f, g1 = 0.f, h00 = 0.f, h01 = 0.f, h11 = 0.f;
        for (int e = e0; e < e1; ++e) {
            const int ent = pv.node_ent[e];
            const float* b = tsum + (int64_t)(ent >> 2) * 20 + 5 * (ent & 3);
            g0 += b[0];
            g1 += b[1];
            h00 += b[2];
            h01 += b[3];
            h11 += b[4];
        }
        s2 = lgn_node_update(n, u, g0, g1, h00, h01, h11, e1 > e0, nbi, nbd, nbw, nrw, mdn, lam, max_step, alpha,
                             omega, u_next);
    }
    block_sum_store(s2, s2_part + blockIdx.x);
}

// hist[k] = (sqrt(sum r^2 / Np), sqrt(sum step^2 / (2 Nn))), partials summed in order.
__global__ void k_lgn_hist(const double* __restrict__ r2_part, int64_t nr, const double* __restrict__ s2_part,
                           int64_t ns, int64_t np, int64_t nn, float* __restrict__ hist_k) {
    __shared__ double s[2][kLgnBlock];
    double a = 0.0, b = 0.0;
    for (int64_t i = threadIdx.x; i < nr; i += blockDim.x) a += r2_part[i];
    for (int64_t i = threadIdx.x; i < ns; i += blockDim.x) b += s2_part[i];
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0.0, y = 0.0;
        for (int i = 0; i < kLgnBlock; ++i) {
            x += s[0][i];
            y += s[1][i];
        }
        hist_k[0] = (float)sqrt(x / (double)std::max<int64_t>(np, 1));
        hist_k[1] = (float)sqrt(y / (double)std::max<int64_t>(2 * nn, 1));
    }
}

}  // namespace

extern "C" {

int d2ic_lgn_solve(const float* pixel_coords, const int32_t* pixel_nodes, const float* pixel_shapeN, int64_t np,
                   const int32_t* node_pixel_index, const float* node_N_weight, const int32_t* node_pixel_degree,
                   int64_t max_deg_pix, const int32_t* nb_idx, const int32_t* nb_deg, const float* nb_w,
                   const float* node_reg_weight, int64_t max_deg_nb, int64_t nn, const float* ref_image,
                   const float* def_image, int32_t H, int32_t W, int32_t interp, float* u, int32_t n_sweeps, float lam,
                   float max_step, float alpha_reg, float omega, float* hist, void* stream) {
    D2IC_CHECK_ARG(u && ref_image && def_image, "u and the images must not be NULL");
    D2IC_CHECK_ARG(np >= 0 && nn >= 0 && H > 0 && W > 0 && n_sweeps >= 0, "sizes must be non-negative");
    D2IC_CHECK_ARG(interp == 0 || interp == 1, "interp must be 0 (cubic) or 1 (linear)");
    D2IC_CHECK_ARG(np == 0 || (pixel_coords && pixel_nodes && pixel_shapeN), "pixel arrays must not be NULL");
    D2IC_CHECK_ARG(nn == 0 || (node_pixel_degree && nb_deg && node_reg_weight), "node arrays must not be NULL");
    D2IC_CHECK_ARG(max_deg_pix == 0 || (node_pixel_index && node_N_weight), "node->pixel table must not be NULL");
    D2IC_CHECK_ARG(max_deg_nb == 0 || (nb_idx && nb_w), "neighbour table must not be NULL");
    hipStream_t st = (hipStream_t)stream;
    if (n_sweeps == 0 || nn == 0) return D2IC_OK;
    const int64_t hw = (int64_t)H * W;
    const int64_t bp = std::max<int64_t>(ceil_div(np, kLgnBlock), 1), bn = ceil_div(nn, kLgnNodesPerBlock);
    const size_t bytes = sizeof(float) * (2 * hw + 4 * std::max<int64_t>(np, 1) + 2 * nn) +
                         sizeof(double) * (bp + bn) + 256;
    void* mem = nullptr;
    D2IC_HIP(hipMallocAsync(&mem, bytes, st));
    char* cur = reinterpret_cast<char*>(mem);
    auto take = [&](size_t b) {
        char* p = cur;
        cur += (b + 15) / 16 * 16;
        return p;
    };
    double* r2_part = reinterpret_cast<double*>(take(sizeof(double) * bp));
    double* s2_part = reinterpret_cast<double*>(take(sizeof(double) * bn));
    float* gximg = reinterpret_cast<float*>(take(sizeof(float) * hw));
    float* gyimg = reinterpret_cast<float*>(take(sizeof(float) * hw));
    float* I1 = reinterpret_cast<float*>(take(sizeof(float) * std::max<int64_t>(np, 1)));
    float* r = reinterpret_cast<float*>(take(sizeof(float) * std::max<int64_t>(np, 1)));
    float* gxd = reinterpret_cast<float*>(take(sizeof(float) * std::max<int64_t>(np, 1)));
    float* gyd = reinterpret_cast<float*>(take(sizeof(float) * std::max<int64_t>(np, 1)));
    float2* ub = reinterpret_cast<float2*>(take(sizeof(float) * 2 * nn));
    k_lgn_grad<<<(unsigned)ceil_div(hw, 256), 256, 0, st>>>(def_image, H, W, gximg, gyimg);
    D2IC_LAUNCH_CHECK();
    if (np > 0) {
        k_lgn_ref<<<(unsigned)bp, kLgnBlock, 0, st>>>(reinterpret_cast<const float2*>(pixel_coords), np, ref_image, H, W,
                                                      interp, I1);
        D2IC_LAUNCH_CHECK();
    }
    float2* ucur = reinterpret_cast<float2*>(u);
    float2* unext = ub;
    for (int k = 0; k < n_sweeps; ++k) {
        k_lgn_pixels<<<(unsigned)bp, kLgnBlock, 0, st>>>(
            reinterpret_cast<const float2*>(pixel_coords), reinterpret_cast<const int4*>(pixel_nodes),
            reinterpret_cast<const float4*>(pixel_shapeN), np, ucur, def_image, gximg, gyimg, H, W, interp, I1, r, gxd,
            gyd, r2_part);
        D2IC_LAUNCH_CHECK();
        k_lgn_nodes<<<(unsigned)bn, kLgnBlock, 0, st>>>(ucur, r, gxd, gyd, node_pixel_index, node_N_weight,
                                                         node_pixel_degree, max_deg_pix, nb_idx, nb_deg, nb_w,
                                                         node_reg_weight, max_deg_nb, nn, lam, max_step, alpha_reg,
                                                         omega, unext, s2_part);
        D2IC_LAUNCH_CHECK();
        if (hist) {
            k_lgn_hist<<<1, kLgnBlock, 0, st>>>(r2_part, bp, s2_part, bn, np, nn, hist + 2 * (size_t)k);
            D2IC_LAUNCH_CHECK();
        }
        std::swap(ucur, unext);
    }
    if (ucur != reinterpret_cast<float2*>(u))
        D2IC_HIP(hipMemcpyAsync(u, ucur, sizeof(float) * 2 * nn, hipMemcpyDeviceToDevice, st));
    D2IC_HIP(hipFreeAsync(mem, st));
    return D2IC_OK;
}

int d2ic_lgn_solve_plan(const d2ic_cg_plan* plan, const int32_t* nb_idx, const int32_t* nb_deg, const float* nb_w,
                        const float* node_reg_weight, int64_t max_deg_nb, const float* ref_image,
                        const float* def_image, int32_t H, int32_t W, int32_t interp, float* u, int32_t n_sweeps,
                        float lam, float max_step, float alpha_reg, float omega, float* hist, void* stream) {
    D2IC_CHECK_ARG(plan && u && ref_image && def_image, "plan, u and the images must not be NULL");
    D2IC_CHECK_ARG(H > 0 && W > 0 && n_sweeps >= 0, "sizes must be non-negative");
    D2IC_CHECK_ARG(interp == 0 || interp == 1, "interp must be 0 (cubic) or 1 (linear)");
    const int64_t nn = plan->nn, nt = plan->ntiles;
    D2IC_CHECK_ARG(nn == 0 || (nb_deg && node_reg_weight), "node arrays must not be NULL");
    D2IC_CHECK_ARG(max_deg_nb == 0 || (nb_idx && nb_w), "neighbour table must not be NULL");
    hipStream_t st = (hipStream_t)stream;
    if (n_sweeps == 0 || nn == 0) return D2IC_OK;
    const PlanView pv = plan_view(plan);
    const int64_t hw = (int64_t)H * W, bn = ceil_div(nn, kLgnBlock), nslot = std::max<int64_t>(nt * kTile, 1);
    const size_t bytes = sizeof(float) * (2 * hw + nslot + 20 * std::max<int64_t>(nt, 1) + 2 * nn) +
                         sizeof(double) * (std::max<int64_t>(nt, 1) + bn) + 256;
    void* mem = nullptr;
    D2IC_HIP(hipMallocAsync(&mem, bytes, st));
    char* cur = reinterpret_cast<char*>(mem);
    auto take = [&](size_t b) {
        char* p = cur;
        cur += (b + 15) / 16 * 16;
        return p;
    };
    double* r2_part = reinterpret_cast<double*>(take(sizeof(double) * std::max<int64_t>(nt, 1)));
    double* s2_part = reinterpret_cast<double*>(take(sizeof(double) * bn));
    float* gximg = reinterpret_cast<float*>(take(sizeof(float) * hw));
    float* gyimg = reinterpret_cast<float*>(take(sizeof(float) * hw));
    float* I1 = reinterpret_cast<float*>(take(sizeof(float) * nslot));
    float* tsum = reinterpret_cast<float*>(take(sizeof(float) * 20 * std::max<int64_t>(nt, 1)));
    float2* ub = reinterpret_cast<float2*>(take(sizeof(float) * 2 * nn));
    k_lgn_grad<<<(unsigned)ceil_div(hw, 256), 256, 0, st>>>(def_image, H, W, gximg, gyimg);
    D2IC_LAUNCH_CHECK();
    if (nt > 0) {
        k_lgn_ref_tiles<<<(unsigned)nt, kLgnBlock, 0, st>>>(pv, ref_image, H, W, interp, I1);
        D2IC_LAUNCH_CHECK();
    }
    float2* ucur = reinterpret_cast<float2*>(u);
    float2* unext = ub;
    for (int k = 0; k < n_sweeps; ++k) {
        if (nt > 0) {
            k_lgn_tiles<<<(unsigned)nt, kLgnBlock, 0, st>>>(pv, ucur, def_image, gximg, gyimg, H, W, interp, I1, tsum,
                                                            r2_part);
            D2IC_LAUNCH_CHECK();
        }
        k_lgn_nodes_tiles<<<(unsigned)bn, kLgnBlock, 0, st>>>(pv, ucur, tsum, nb_idx, nb_deg, nb_w, node_reg_weight,
                                                               max_deg_nb, nn, lam, max_step, alpha_reg, omega, unext,
                                                               s2_part);
        D2IC_LAUNCH_CHECK();
        if (hist) {
            k_lgn_hist<<<1, kLgnBlock, 0, st>>>(r2_part, nt, s2_part, bn, plan->np, nn, hist + 2 * (size_t)k);
            D2IC_LAUNCH_CHECK();
        }
        std::swap(ucur, unext);
    }
    if (ucur != reinterpret_cast<float2*>(u))
        D2IC_HIP(hipMemcpyAsync(u, ucur, sizeof(float) * 2 * nn, hipMemcpyDeviceToDevice, st));
    D2IC_HIP(hipFreeAsync(mem, st));
    return D2IC_OK;
}

}  // extern "C"

}  // namespace d2ic
