// Implicit-function derivatives of the global solve.
//
// At a solution u* of grad_u J(u; f, g, alpha) = 0 with
//   J = 1/2 mean_p (g(x_p + u_p) - f_p)^2 + alpha E_spring(u)
// (reference objective: d2ic/solver_global_cg.py:205-323), a loss L(u*) has
//   dL/dtheta = -(dG/dtheta)^T lambda,  H lambda = dL/du,
// with G = grad_u J and H = dG/du.  The kernels here provide, on the
// solver's own tile layout (plan.hpp):
//   * the linearisation at u: per pixel r, dg/dx, dg/dy and per tile the
//     8 x 8 element Hessian sum_p N_a N_b (grad g grad g^T + r hess g)
//     (Gauss-Newton: without the r hess g term);
//   * Hessian-vector products for many vectors (element matvec per tile,
//     deterministic node gather, spring stencil);
//   * tangent right-hand sides -(dG/dg) dg or -(dG/df) df for perturbation
//     images (forward-mode probes);
//   * the transpose: dL/df and dL/dg images for a given lambda.
#include <algorithm>

#include "common.hpp"
#include "interp.hpp"
#include "plan.hpp"

namespace d2ic {

constexpr int kK = 32;  // stored entries per tile: 10 node pairs (a <= b) x 3 (xx, xy, yy), padded

__device__ __forceinline__ int pair_index(int a, int b) {  // a <= b
    return a * 4 - (a * (a - 1)) / 2 + (b - a);
}

__device__ __forceinline__ int node_of(int4 nd, int a) { return a == 0 ? nd.x : (a == 1 ? nd.y : (a == 2 ? nd.z : nd.w)); }

// Sampling position of slot s at displacement u (same arithmetic as the solver).
__device__ __forceinline__ void slot_position(const PlanView& pl, const float* __restrict__ u, int64_t s, int4 nd,
                                              int& x0, float& fx, int& y0, float& fy, float4& N) {
    N = pl.N4[s];
    const float ux = N.x * u[2 * nd.x] + N.y * u[2 * nd.y] + N.z * u[2 * nd.z] + N.w * u[2 * nd.w];
    const float uy = N.x * u[2 * nd.x + 1] + N.y * u[2 * nd.y + 1] + N.z * u[2 * nd.z + 1] + N.w * u[2 * nd.w + 1];
    split_coord(pl.px[s], ux, x0, fx);
    split_coord(pl.py[s], uy, y0, fy);
}

// One block (kTile threads) per tile.
__global__ __launch_bounds__(kTile) void k_adj_lin(PlanView pl, const float* __restrict__ img, int H, int W,
                                                   const float* __restrict__ u, int gauss_newton,
                                                   float* __restrict__ lin, float* __restrict__ K) {
    __shared__ float s_part[kTile / kWave][30];
    const int t = blockIdx.x;
    const int i = threadIdx.x;
    const int lane = i & (kWave - 1), wave = i / kWave;
    const int cnt = pl.tile_cnt[t];
    const int4 nd = pl.tile_nodes[t];
    const int64_t s = (int64_t)t * kTile + i;
    float v[30];
#pragma unroll
    for (int k = 0; k < 30; ++k) v[k] = 0.f;
    float r = 0.f, gx = 0.f, gy = 0.f;
    if (i < cnt) {
        int x0, y0;
        float fx, fy;
        float4 N;
        slot_position(pl, u, s, nd, x0, fx, y0, fy, N);
        float o[6];
        sample_hess(pl.interp, img, H, W, x0, fx, y0, fy, o);
        r = o[0] - pl.F[s];
        gx = o[1];
        gy = o[2];
        const float m00 = gx * gx + (gauss_newton ? 0.f : r * o[3]);
        const float m01 = gx * gy + (gauss_newton ? 0.f : r * o[4]);
        const float m11 = gy * gy + (gauss_newton ? 0.f : r * o[5]);
        const float Na[4] = {N.x, N.y, N.z, N.w};
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = a; b < 4; ++b) {
                const float nn = Na[a] * Na[b];
                const int q = pair_index(a, b);
                v[3 * q] = nn * m00;
                v[3 * q + 1] = nn * m01;
                v[3 * q + 2] = nn * m11;
            }
    }
    lin[3 * s] = r;
    lin[3 * s + 1] = gx;
    lin[3 * s + 2] = gy;
#pragma unroll
    for (int k = 0; k < 30; ++k) {
        const float w = wave_sum(v[k]);
        if (lane == 0) s_part[wave][k] = w;
    }
    __syncthreads();
    if (i < kK) {
        float acc = 0.f;
        if (i < 30)
            for (int w = 0; w < kTile / kWave; ++w) acc += s_part[w][i];
        K[(int64_t)t * kK + i] = acc;
    }
}

// y_t = K_t v_t for every tile and vector (thread per (tile, vector)).
__global__ void k_adj_hv_tiles(PlanView pl, const float* __restrict__ K, const float* __restrict__ V, int nv,
                               float* __restrict__ TPV) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (int64_t)nv * pl.ntiles) return;
    const int k = (int)(id / pl.ntiles);
    const int64_t t = id % pl.ntiles;
    const int4 nd = pl.tile_nodes[t];
    const float* v = V + (size_t)k * pl.nn * 2;
    float vx[4], vy[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int n = node_of(nd, a);
        vx[a] = v[2 * n];
        vy[a] = v[2 * n + 1];
    }
    const float* Kt = K + t * kK;
    float* o = TPV + ((size_t)k * pl.ntiles + t) * 8;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        float yx = 0.f, yy = 0.f;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int q = a <= b ? pair_index(a, b) : pair_index(b, a);
            const float k00 = Kt[3 * q], k01 = Kt[3 * q + 1], k11 = Kt[3 * q + 2];
            yx += k00 * vx[b] + k01 * vy[b];
            yy += k01 * vx[b] + k11 * vy[b];
        }
        o[2 * a] = yx;
        o[2 * a + 1] = yy;
    }
}

// out_k[n] = scale * sum over the node's (tile, slot) entries of TPV_k, in
// table order, + alpha * (L_W v_k)[n] when V is given.
__global__ void k_adj_gather(PlanView pl, const float* __restrict__ TPV, int nv, double scale, float alpha,
                             const float* __restrict__ V, float* __restrict__ out) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (int64_t)nv * pl.nn) return;
    const int k = (int)(id / pl.nn);
    const int n = (int)(id % pl.nn);
    double sx = 0.0, sy = 0.0;
    const float* tp = TPV + (size_t)k * pl.ntiles * 8;
    for (int e = pl.node_ptr[n]; e < pl.node_ptr[n + 1]; ++e) {
        const int ent = pl.node_ent[e];
        sx += (double)tp[2 * (size_t)ent];
        sy += (double)tp[2 * (size_t)ent + 1];
    }
    float rx = 0.f, ry = 0.f;
    if (V) {
        const float* v = V + (size_t)k * pl.nn * 2;
        for (int j = 0; j < pl.nb_maxdeg; ++j) {
            const int m = pl.nb_idx[(size_t)n * pl.nb_maxdeg + j];
            if (m < 0) continue;
            const float w = pl.nb_wp[(size_t)n * pl.nb_maxdeg + j];
            rx += w * (v[2 * n] - v[2 * m]);
            ry += w * (v[2 * n + 1] - v[2 * m + 1]);
        }
    }
    out[(size_t)k * pl.nn * 2 + 2 * n] = (float)(sx * scale) + alpha * rx;
    out[(size_t)k * pl.nn * 2 + 2 * n + 1] = (float)(sy * scale) + alpha * ry;
}

// Inverse 2 x 2 diagonal blocks of H (block-Jacobi preconditioner): for node n,
// B = scale * sum over its (tile, slot a) entries of the tile's (a, a) block
// + alpha * sum_j w_nj * I, in table order; stored as (ixx, ixy, iyy).  A
// block that is not positive definite (the r hess g term can make it so)
// falls back to 1 / (|Bxx| + |Byy|) * I; a zero block gives 0.
__global__ void k_adj_diag_inv(PlanView pl, const float* __restrict__ K, double scale, float alpha,
                               float* __restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= pl.nn) return;
    double bxx = 0.0, bxy = 0.0, byy = 0.0;
    for (int e = pl.node_ptr[n]; e < pl.node_ptr[n + 1]; ++e) {
        const int ent = pl.node_ent[e];
        const int t = ent >> 2, a = ent & 3;
        const float* Kt = K + (int64_t)t * kK + 3 * pair_index(a, a);
        bxx += (double)Kt[0];
        bxy += (double)Kt[1];
        byy += (double)Kt[2];
    }
    double sw = 0.0;
    for (int j = 0; j < pl.nb_maxdeg; ++j) {
        const int m = pl.nb_idx[(size_t)n * pl.nb_maxdeg + j];
        if (m >= 0) sw += (double)pl.nb_wp[(size_t)n * pl.nb_maxdeg + j];
    }
    bxx = bxx * scale + (double)alpha * sw;
    byy = byy * scale + (double)alpha * sw;
    bxy = bxy * scale;
    const double det = bxx * byy - bxy * bxy, tr = fabs(bxx) + fabs(byy);
    float ixx = 0.f, ixy = 0.f, iyy = 0.f;
    if (bxx > 0.0 && det > 1e-12 * tr * tr) {
        ixx = (float)(byy / det);
        ixy = (float)(-bxy / det);
        iyy = (float)(bxx / det);
    } else if (tr > 0.0) {
        ixx = iyy = (float)(1.0 / tr);
    }
    out[3 * (size_t)n] = ixx;
    out[3 * (size_t)n + 1] = ixy;
    out[3 * (size_t)n + 2] = iyy;
}

// Tangent partials per tile and perturbation image: sum_p N_a c_p with
//   which 0 (g): c_p = dg(x_p) grad g(x_p) + r_p grad dg(x_p)
//   which 1 (f): c_p = -df(pixel p) grad g(x_p)
// One block (kTile threads) per tile, looping over the images.
__global__ __launch_bounds__(kTile) void k_adj_tangent_tiles(PlanView pl, int H, int W,
                                                             const float* __restrict__ u,
                                                             const float* __restrict__ lin, int which,
                                                             const float* __restrict__ dimg, int nv,
                                                             float* __restrict__ TPV) {
    __shared__ float s_part[kTile / kWave][8];
    const int t = blockIdx.x;
    const int i = threadIdx.x;
    const int lane = i & (kWave - 1), wave = i / kWave;
    const int cnt = pl.tile_cnt[t];
    const int4 nd = pl.tile_nodes[t];
    const int64_t s = (int64_t)t * kTile + i;
    int x0 = 0, y0 = 0, col = 0, row = 0;
    float fx = 0.f, fy = 0.f;
    float4 N = make_float4(0.f, 0.f, 0.f, 0.f);
    const bool live = i < cnt;
    if (live) {
        slot_position(pl, u, s, nd, x0, fx, y0, fy, N);
        col = min(max((int)floorf(pl.px[s]), 0), W - 1);
        row = min(max((int)floorf(pl.py[s]), 0), H - 1);
    }
    const float r = lin[3 * s], gx = lin[3 * s + 1], gy = lin[3 * s + 2];
    for (int k = 0; k < nv; ++k) {
        const float* im = dimg + (size_t)k * H * W;
        float cx = 0.f, cy = 0.f;
        if (live) {
            if (which == 0) {
                float dv, dgx, dgy;
                sample_at(pl.interp, im, H, W, x0, fx, y0, fy, dv, dgx, dgy);
                cx = dv * gx + r * dgx;
                cy = dv * gy + r * dgy;
            } else {
                const float df = im[(size_t)row * W + col];
                cx = -df * gx;
                cy = -df * gy;
            }
        }
        const float c8[8] = {N.x * cx, N.x * cy, N.y * cx, N.y * cy, N.z * cx, N.z * cy, N.w * cx, N.w * cy};
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float w = wave_sum(c8[q]);
            if (lane == 0) s_part[wave][q] = w;
        }
        __syncthreads();
        if (i < 8) {
            float acc = 0.f;
            for (int w = 0; w < kTile / kWave; ++w) acc += s_part[w][i];
            TPV[((size_t)k * pl.ntiles + t) * 8 + i] = acc;
        }
        __syncthreads();
    }
}

constexpr double kAdjFix = 1099511627776.0;  // 2^40: fixed-point scale of the image-gradient accumulator

__device__ __forceinline__ void fix_add(unsigned long long* acc, double v) {
    atomicAdd(acc, (unsigned long long)(long long)llrint(v * kAdjFix));
}

// Transpose of the image terms for a given lambda (thread per pixel slot):
//   dL/dg[q] -= inv_np (w_q(x_p) s_p + r_p lambda_p . grad w_q(x_p))   (16 or 4 taps)
//   dL/df[pixel p] = inv_np s_p,   s_p = lambda_p . grad g(x_p).
// The g-image sums are exact fixed-point integer sums (order independent).
__global__ void k_adj_image(PlanView pl, int H, int W, const float* __restrict__ u,
                            const float* __restrict__ lin, const float* __restrict__ lam, double inv_np,
                            float* __restrict__ grad_f, unsigned long long* __restrict__ acc_g) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= pl.ntiles * kTile) return;
    const int64_t t = s / kTile;
    if ((int)(s % kTile) >= pl.tile_cnt[t]) return;
    const int4 nd = pl.tile_nodes[t];
    int x0, y0;
    float fx, fy;
    float4 N;
    slot_position(pl, u, s, nd, x0, fx, y0, fy, N);
    const float lx = N.x * lam[2 * nd.x] + N.y * lam[2 * nd.y] + N.z * lam[2 * nd.z] + N.w * lam[2 * nd.w];
    const float ly =
        N.x * lam[2 * nd.x + 1] + N.y * lam[2 * nd.y + 1] + N.z * lam[2 * nd.z + 1] + N.w * lam[2 * nd.w + 1];
    const float r = lin[3 * s], gx = lin[3 * s + 1], gy = lin[3 * s + 2];
    const double sp = (double)lx * gx + (double)ly * gy;
    const int col = min(max((int)floorf(pl.px[s]), 0), W - 1);
    const int row = min(max((int)floorf(pl.py[s]), 0), H - 1);
    if (grad_f) grad_f[(size_t)row * W + col] = (float)(inv_np * sp);
    if (!acc_g) return;
    int cols[4], rows[4];
    float wx[4], dwx[4], wy[4], dwy[4];
    const int n = sampler_taps(pl.interp, H, W, x0, fx, y0, fy, cols, rows, wx, dwx, wy, dwy);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const double w = (double)wx[j] * wy[i];
            const double dw = (double)lx * ((double)dwx[j] * wy[i]) + (double)ly * ((double)wx[j] * dwy[i]);
            fix_add(&acc_g[(size_t)rows[i] * W + cols[j]], -inv_np * (w * sp + (double)r * dw));
        }
}

__global__ void k_fix_to_float(const unsigned long long* __restrict__ acc, int64_t n, float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)((double)(long long)acc[i] / kAdjFix);
}

__global__ void k_zero_u64(unsigned long long* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

}  // namespace d2ic

using namespace d2ic;

extern "C" {

int d2ic_adj_sizes(const d2ic_cg_plan* plan, int64_t* lin_floats, int64_t* k_floats, int64_t* tile_floats) {
    D2IC_CHECK_ARG(plan != nullptr, "plan must not be NULL");
    if (lin_floats) *lin_floats = 3 * plan->ntiles * kTile;
    if (k_floats) *k_floats = kK * plan->ntiles;
    if (tile_floats) *tile_floats = 8 * plan->ntiles;
    return D2IC_OK;
}

int d2ic_adj_linearize(const d2ic_cg_plan* plan, const float* def_image, int32_t H, int32_t W, const float* u,
                       int32_t gauss_newton, float* lin, float* K, void* stream) {
    D2IC_CHECK_ARG(plan && def_image && u && lin && K, "arguments must not be NULL");
    D2IC_CHECK_ARG(H > 0 && W > 0, "image must be non-empty");
    D2IC_CHECK_ARG(plan->has_ref, "plan has no reference image");
    if (plan->ntiles == 0) return D2IC_OK;
    hipStream_t st = (hipStream_t)stream;
    k_adj_lin<<<(unsigned)plan->ntiles, kTile, 0, st>>>(plan_view(plan), def_image, H, W, u, gauss_newton, lin, K);
    D2IC_LAUNCH_CHECK();
    return D2IC_OK;
}

int d2ic_adj_hvp(const d2ic_cg_plan* plan, const float* K, float alpha_reg, const float* V, int32_t nv, float* out,
                 float* work, void* stream) {
    D2IC_CHECK_ARG(plan && K && V && out && work, "arguments must not be NULL");
    D2IC_CHECK_ARG(nv >= 1, "nv must be >= 1");
    hipStream_t st = (hipStream_t)stream;
    const int64_t nt = (int64_t)nv * plan->ntiles, nn = (int64_t)nv * plan->nn;
    if (nt > 0) {
        k_adj_hv_tiles<<<ceil_div(nt, 256), 256, 0, st>>>(plan_view(plan), K, V, nv, work);
        D2IC_LAUNCH_CHECK();
    }
    if (nn > 0) {
        k_adj_gather<<<ceil_div(nn, 256), 256, 0, st>>>(plan_view(plan), work, nv, 1.0 / (double)std::max<int64_t>(plan->np, 1),
                                                         alpha_reg, V, out);
        D2IC_LAUNCH_CHECK();
    }
    return D2IC_OK;
}

int d2ic_adj_tangent(const d2ic_cg_plan* plan, int32_t H, int32_t W, const float* u, const float* lin, int32_t which,
                     const float* dimg, int32_t nv, float* b, float* work, void* stream) {
    D2IC_CHECK_ARG(plan && u && lin && dimg && b && work, "arguments must not be NULL");
    D2IC_CHECK_ARG(which == 0 || which == 1, "which must be 0 (deformed image) or 1 (reference image)");
    D2IC_CHECK_ARG(nv >= 1 && H > 0 && W > 0, "nv and the image size must be positive");
    hipStream_t st = (hipStream_t)stream;
    if (plan->ntiles > 0) {
        k_adj_tangent_tiles<<<(unsigned)plan->ntiles, kTile, 0, st>>>(plan_view(plan), H, W, u, lin, which, dimg, nv, work);
        D2IC_LAUNCH_CHECK();
    }
    const int64_t nn = (int64_t)nv * plan->nn;
    if (nn > 0) {
        k_adj_gather<<<ceil_div(nn, 256), 256, 0, st>>>(
            plan_view(plan), work, nv, -1.0 / (double)std::max<int64_t>(plan->np, 1), 0.f, nullptr, b);
        D2IC_LAUNCH_CHECK();
    }
    return D2IC_OK;
}

int d2ic_adj_image_grads(const d2ic_cg_plan* plan, int32_t H, int32_t W, const float* u, const float* lin,
                         const float* lam, float* grad_f, float* grad_g, unsigned long long* work, void* stream) {
    D2IC_CHECK_ARG(plan && u && lin && lam, "arguments must not be NULL");
    D2IC_CHECK_ARG(!grad_g || work, "grad_g needs a work buffer of H*W 64-bit words");
    D2IC_CHECK_ARG(H > 0 && W > 0, "image must be non-empty");
    hipStream_t st = (hipStream_t)stream;
    const int64_t npx = (int64_t)H * W;
    if (grad_f) D2IC_HIP(hipMemsetAsync(grad_f, 0, sizeof(float) * npx, st));
    if (grad_g) {
        k_zero_u64<<<ceil_div(npx, 256), 256, 0, st>>>(work, npx);
        D2IC_LAUNCH_CHECK();
    }
    const int64_t ns = plan->ntiles * kTile;
    if (ns > 0) {
        k_adj_image<<<ceil_div(ns, 256), 256, 0, st>>>(plan_view(plan), H, W, u, lin, lam,
                                                        1.0 / (double)std::max<int64_t>(plan->np, 1), grad_f,
                                                        grad_g ? work : nullptr);
        D2IC_LAUNCH_CHECK();
    }
    if (grad_g) {
        k_fix_to_float<<<ceil_div(npx, 256), 256, 0, st>>>(work, npx, grad_g);
        D2IC_LAUNCH_CHECK();
    }
    return D2IC_OK;
}

int d2ic_adj_diag_inv(const d2ic_cg_plan* plan, const float* K, float alpha_reg, float* out, void* stream) {
    D2IC_CHECK_ARG(plan && K && out, "arguments must not be NULL");
    hipStream_t st = (hipStream_t)stream;
    if (plan->nn > 0) {
        k_adj_diag_inv<<<ceil_div(plan->nn, 256), 256, 0, st>>>(plan_view(plan), K,
                                                               1.0 / (double)std::max<int64_t>(plan->np, 1), alpha_reg, out);
        D2IC_LAUNCH_CHECK();
    }
    return D2IC_OK;
}

int d2ic_adj_spring(const d2ic_cg_plan* plan, const float* V, int32_t nv, float* out, float* work, void* stream) {
    D2IC_CHECK_ARG(plan && V && out && work, "arguments must not be NULL");
    D2IC_CHECK_ARG(nv >= 1, "nv must be >= 1");
    hipStream_t st = (hipStream_t)stream;
    const int64_t nt = (int64_t)nv * plan->ntiles * 8;
    if (nt > 0) D2IC_HIP(hipMemsetAsync(work, 0, sizeof(float) * nt, st));
    const int64_t nn = (int64_t)nv * plan->nn;
    if (nn > 0) {
        k_adj_gather<<<ceil_div(nn, 256), 256, 0, st>>>(plan_view(plan), work, nv, 0.0, 1.0f, V, out);
        D2IC_LAUNCH_CHECK();
    }
    return D2IC_OK;
}

}  // extern "C"
