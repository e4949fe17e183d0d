// Pixel assets on the device: pixel->element first match, Q4 shape functions
// by Newton, node->pixel table (reference: d2ic/pixel_assets.py:21-526).
//
// Built with -ffp-contract=off: every product/sum is rounded as written so the
// integer results (element index, slot order) and the Newton iterates follow
// the reference's IEEE float32 sequence.
#include <hipcub/hipcub.hpp>

#include <cstdarg>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "interp.hpp"

namespace d2ic {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

const char* last_error() { return g_last_error.c_str(); }

// ---------------------------------------------------------------------------
// Quad geometry (pixel_assets.py:169-187): corners, edges, inclusive bbox and
// the orientation sign of the shoelace area (0 counts as positive).
// ---------------------------------------------------------------------------
struct Quad {
    float qx[4], qy[4], ex[4], ey[4];
    float bx0, bx1, by0, by1;
    float orient;
    float pad[3];
};

__global__ void k_quads(const float* __restrict__ nodes, const int32_t* __restrict__ elems, int64_t ne,
                        Quad* __restrict__ quads) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    Quad q;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int n = elems[4 * e + k];
        q.qx[k] = nodes[2 * (int64_t)n];
        q.qy[k] = nodes[2 * (int64_t)n + 1];
    }
    q.bx0 = q.bx1 = q.qx[0];
    q.by0 = q.by1 = q.qy[0];
#pragma unroll
    for (int k = 1; k < 4; ++k) {
        q.bx0 = q.qx[k] < q.bx0 ? q.qx[k] : q.bx0;
        q.bx1 = q.qx[k] > q.bx1 ? q.qx[k] : q.bx1;
        q.by0 = q.qy[k] < q.by0 ? q.qy[k] : q.by0;
        q.by1 = q.qy[k] > q.by1 ? q.qy[k] : q.by1;
    }
    float area2 = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int kn = (k + 1) & 3;
        q.ex[k] = q.qx[kn] - q.qx[k];
        q.ey[k] = q.qy[kn] - q.qy[k];
        area2 = area2 + (q.qx[k] * q.qy[kn] - q.qy[k] * q.qx[kn]);
    }
    q.orient = area2 >= 0.f ? 1.f : -1.f;
    q.pad[0] = q.pad[1] = q.pad[2] = 0.f;
    quads[e] = q;
}

// _assign_chunk's inside test (pixel_assets.py:189-202).
__device__ __forceinline__ bool quad_contains(const Quad& q, float x, float y) {
    if (!(x >= q.bx0 && x <= q.bx1 && y >= q.by0 && y <= q.by1)) return false;
    const float tol = 1e-6f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float vx = x - q.qx[k], vy = y - q.qy[k];
        float cr = q.ex[k] * vy - q.ey[k] * vx;
        if (!(q.orient * cr >= -tol)) return false;
    }
    return true;
}

// Uniform bin grid over element bounding boxes.  An element is registered in
// every bin its inclusive bbox overlaps; a pixel inside the bbox falls in one
// of those bins because floor((x - x0)/bs) is monotone.  The first match is
// the minimum passing element index among the pixel's bin entries, which is
// the reference's lowest passing index.
struct BinGrid {
    double x0, y0, bs;
    int nbx, nby;
};

__device__ __forceinline__ void bin_range(const BinGrid& g, const Quad& q, int& bx0, int& bx1, int& by0, int& by1) {
    bx0 = (int)floor(((double)q.bx0 - g.x0) / g.bs);
    bx1 = (int)floor(((double)q.bx1 - g.x0) / g.bs);
    by0 = (int)floor(((double)q.by0 - g.y0) / g.bs);
    by1 = (int)floor(((double)q.by1 - g.y0) / g.bs);
    bx0 = bx0 < 0 ? 0 : bx0;
    by0 = by0 < 0 ? 0 : by0;
    bx1 = bx1 > g.nbx - 1 ? g.nbx - 1 : bx1;
    by1 = by1 > g.nby - 1 ? g.nby - 1 : by1;
}

__global__ void k_bin_count(const Quad* __restrict__ quads, int64_t ne, BinGrid g, int32_t* __restrict__ counts) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    int bx0, bx1, by0, by1;
    bin_range(g, quads[e], bx0, bx1, by0, by1);
    for (int by = by0; by <= by1; ++by)
        for (int bx = bx0; bx <= bx1; ++bx) atomicAdd(&counts[by * g.nbx + bx], 1);
}

__global__ void k_bin_fill(const Quad* __restrict__ quads, int64_t ne, BinGrid g, const int32_t* __restrict__ start,
                           int32_t* __restrict__ fill, int32_t* __restrict__ list) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    int bx0, bx1, by0, by1;
    bin_range(g, quads[e], bx0, bx1, by0, by1);
    for (int by = by0; by <= by1; ++by)
        for (int bx = bx0; bx <= bx1; ++bx) {
            int b = by * g.nbx + bx;
            int pos = atomicAdd(&fill[b], 1);
            list[start[b] + pos] = (int32_t)e;
        }
}

// One thread per bbox candidate (row-major, pixel_assets.py:115-132): ROI
// filter then first-match search.
__global__ void k_assign(const Quad* __restrict__ quads, BinGrid g, const int32_t* __restrict__ bstart,
                         const int32_t* __restrict__ blist, int32_t H, int32_t W, int32_t i0, int32_t j0, int32_t nj,
                         int64_t ncand, const uint8_t* __restrict__ roi, int32_t* __restrict__ cand_elt) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncand) return;
    const int i = i0 + (int)(t / nj), j = j0 + (int)(t % nj);
    if (roi) {
        // roi_rows = clip(floor(y - 0.5)), roi_cols = clip(floor(x - 0.5)) (pixel_assets.py:70-72)
        int rr = clampi(i, 0, H - 1), cc = clampi(j, 0, W - 1);
        if (!roi[(int64_t)rr * W + cc]) {
            cand_elt[t] = -1;
            return;
        }
    }
    const float x = (float)j + 0.5f, y = (float)i + 0.5f;
    int32_t found = -1;
    double fx = ((double)x - g.x0) / g.bs, fy = ((double)y - g.y0) / g.bs;
    if (fx >= 0.0 && fy >= 0.0) {
        int bx = (int)floor(fx), by = (int)floor(fy);
        if (bx < g.nbx && by < g.nby) {
            int b = by * g.nbx + bx;
            for (int k = bstart[b]; k < bstart[b + 1]; ++k) {
                int32_t e = blist[k];
                if ((found < 0 || e < found) && quad_contains(quads[e], x, y)) found = e;
            }
        }
    }
    cand_elt[t] = found;
}

__global__ void k_node_bbox(const float* __restrict__ nodes, int64_t nn, float* __restrict__ out) {
    // single block: min/max of node coordinates
    __shared__ float s[4][256];
    float mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
    for (int64_t k = threadIdx.x; k < nn; k += blockDim.x) {
        float x = nodes[2 * k], y = nodes[2 * k + 1];
        mnx = fminf(mnx, x); mxx = fmaxf(mxx, x);
        mny = fminf(mny, y); mxy = fmaxf(mxy, y);
    }
    s[0][threadIdx.x] = mnx; s[1][threadIdx.x] = mny; s[2][threadIdx.x] = mxx; s[3][threadIdx.x] = mxy;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            s[0][threadIdx.x] = fminf(s[0][threadIdx.x], s[0][threadIdx.x + off]);
            s[1][threadIdx.x] = fminf(s[1][threadIdx.x], s[1][threadIdx.x + off]);
            s[2][threadIdx.x] = fmaxf(s[2][threadIdx.x], s[2][threadIdx.x + off]);
            s[3][threadIdx.x] = fmaxf(s[3][threadIdx.x], s[3][threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) out[threadIdx.x] = s[threadIdx.x][0];
}

// ---------------------------------------------------------------------------
// Stable compaction of kept candidates (row-major order preserved).
// ---------------------------------------------------------------------------
constexpr int kCompactBlock = 1024;

__global__ __launch_bounds__(kCompactBlock) void k_count_kept(const int32_t* __restrict__ cand_elt, int64_t n,
                                                               int32_t* __restrict__ blk_cnt) {
    int64_t t = (int64_t)blockIdx.x * kCompactBlock + threadIdx.x;
    int flag = (t < n) && cand_elt[t] >= 0;
    int c = __syncthreads_count(flag);
    if (threadIdx.x == 0) blk_cnt[blockIdx.x] = c;
}

__global__ __launch_bounds__(kCompactBlock) void k_scatter_kept(const int32_t* __restrict__ cand_elt, int64_t n,
                                                                 const int32_t* __restrict__ blk_off,
                                                                 int32_t* __restrict__ kept_src) {
    __shared__ int wave_off[kCompactBlock / kWave];
    int64_t t = (int64_t)blockIdx.x * kCompactBlock + threadIdx.x;
    int flag = (t < n) && cand_elt[t] >= 0;
    uint64_t ballot = __ballot(flag);
    int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    int within = __popcll(ballot & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    if (lane == 0) wave_off[wv] = __popcll(ballot);
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int k = 0; k < kCompactBlock / kWave; ++k) {
            int c = wave_off[k];
            wave_off[k] = acc;
            acc += c;
        }
    }
    __syncthreads();
    if (flag) kept_src[blk_off[blockIdx.x] + wave_off[wv] + within] = (int32_t)t;
}

// ---------------------------------------------------------------------------
// Q4 shape functions by 10 Newton steps (pixel_assets.py:277-327) with the
// 2x2 solve of jnp.linalg.solve (partial pivoting LU + two triangular solves).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void shape_q4(float xi, float eta, float N[4]) {
    N[0] = 0.25f * (1.f - xi) * (1.f - eta);
    N[1] = 0.25f * (1.f + xi) * (1.f - eta);
    N[2] = 0.25f * (1.f + xi) * (1.f + eta);
    N[3] = 0.25f * (1.f - xi) * (1.f + eta);
}

__device__ __forceinline__ void solve2(float a, float b, float c, float d, float f0, float f1, float& x0, float& x1) {
    if (fabsf(c) > fabsf(a)) {
        float t;
        t = a; a = c; c = t;
        t = b; b = d; d = t;
        t = f0; f0 = f1; f1 = t;
    }
    float l = c / a;
    float u22 = d - l * b;
    float y1 = f1 - l * f0;
    x1 = y1 / u22;
    x0 = (f0 - b * x1) / a;
}

__device__ void newton_q4(float px, float py, const float X[4], const float Y[4], float& xi_o, float& eta_o) {
    float xi = 0.f, eta = 0.f;
    for (int it = 0; it < 10; ++it) {
        float N[4];
        shape_q4(xi, eta, N);
        const float dxi[4] = {-0.25f * (1.f - eta), 0.25f * (1.f - eta), 0.25f * (1.f + eta), -0.25f * (1.f + eta)};
        const float deta[4] = {-0.25f * (1.f - xi), -0.25f * (1.f + xi), 0.25f * (1.f + xi), 0.25f * (1.f - xi)};
        float x = 0.f, y = 0.f, j00 = 0.f, j01 = 0.f, j10 = 0.f, j11 = 0.f;
        for (int a = 0; a < 4; ++a) {
            x = x + N[a] * X[a];
            y = y + N[a] * Y[a];
        }
        for (int a = 0; a < 4; ++a) {
            j00 = j00 + dxi[a] * X[a];
            j10 = j10 + dxi[a] * Y[a];
        }
        for (int a = 0; a < 4; ++a) {
            j01 = j01 + deta[a] * X[a];
            j11 = j11 + deta[a] * Y[a];
        }
        float d0, d1;
        solve2(j00, j01, j10, j11, x - px, y - py, d0, d1);
        xi = xi - d0;
        eta = eta - d1;
    }
    xi_o = xi;
    eta_o = eta;
}

__global__ void k_fill_pixels(const int32_t* __restrict__ kept_src, const int32_t* __restrict__ cand_elt,
                              int64_t np, int32_t H, int32_t W, int32_t i0, int32_t j0, int32_t nj,
                              const float* __restrict__ nodes, const int32_t* __restrict__ elems,
                              float* __restrict__ coords, int32_t* __restrict__ pnodes, float* __restrict__ shapeN,
                              int32_t* __restrict__ pelts, int32_t* __restrict__ roi_flat) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    int32_t t = kept_src[k];
    int i = i0 + t / nj, j = j0 + t % nj;
    int32_t e = cand_elt[t];
    float x = (float)j + 0.5f, y = (float)i + 0.5f;
    coords[2 * k] = x;
    coords[2 * k + 1] = y;
    float X[4], Y[4];
    int32_t nd[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        nd[a] = elems[4 * (int64_t)e + a];
        X[a] = nodes[2 * (int64_t)nd[a]];
        Y[a] = nodes[2 * (int64_t)nd[a] + 1];
        pnodes[4 * k + a] = nd[a];
    }
    pelts[k] = e;
    roi_flat[k] = clampi(i, 0, H - 1) * W + clampi(j, 0, W - 1);
    float xi, eta, N[4];
    newton_q4(x, y, X, Y, xi, eta);
    shape_q4(xi, eta, N);
#pragma unroll
    for (int a = 0; a < 4; ++a) shapeN[4 * k + a] = N[a];
}

// ---------------------------------------------------------------------------
// node->pixel table
// ---------------------------------------------------------------------------
__global__ void k_degree(const int32_t* __restrict__ pnodes, int64_t n4, int32_t* __restrict__ degree) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n4) atomicAdd(&degree[pnodes[t]], 1);
}

__global__ void k_iota(int32_t* __restrict__ v, int64_t n) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) v[t] = (int32_t)t;
}

__global__ void k_table_fill(const int32_t* __restrict__ keys, const int32_t* __restrict__ vals, int64_t n4,
                             const int32_t* __restrict__ node_start, const float* __restrict__ shapeN,
                             int32_t maxdeg, int32_t* __restrict__ idx, float* __restrict__ w) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n4) return;
    int32_t n = keys[s], t = vals[s];
    int64_t slot = s - node_start[n];
    idx[(int64_t)n * maxdeg + slot] = t / 4;
    w[(int64_t)n * maxdeg + slot] = shapeN[t];
}

static int bits_for(int64_t n) {
    int b = 1;
    while ((int64_t(1) << b) < n) ++b;
    return b;
}

}  // namespace d2ic

using namespace d2ic;

// A small RAII holder for stream-ordered temporaries.
struct DevBuf {
    void* p = nullptr;
    hipStream_t s = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFreeAsync(p, s);
    }
    hipError_t alloc(size_t bytes, hipStream_t st) {
        s = st;
        return hipMallocAsync(&p, bytes ? bytes : 16, st);
    }
    template <class T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

extern "C" {

const char* d2ic_last_error(void) { return d2ic::last_error(); }
int d2ic_version(void) { return 100; }

int d2ic_pixel_map(const float* nodes, int64_t n_nodes, const int32_t* elements, int64_t n_elems, int32_t H,
                   int32_t W, int32_t i0, int32_t i1, int32_t j0, int32_t j1, const uint8_t* roi, int32_t* cand_elt,
                   int64_t* n_kept, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    D2IC_CHECK_ARG(H > 0 && W > 0, "image shape must be positive");
    D2IC_CHECK_ARG(n_kept != nullptr, "n_kept must not be NULL");
    D2IC_CHECK_ARG(0 <= i0 && i0 <= i1 && i1 <= H && 0 <= j0 && j0 <= j1 && j1 <= W, "bbox outside image");
    int64_t nj = j1 - j0, ncand = (int64_t)(i1 - i0) * nj;
    D2IC_CHECK_ARG(ncand < (int64_t)INT32_MAX, "too many candidate pixels");
    D2IC_CHECK_ARG(n_nodes >= 0 && n_elems >= 0, "sizes must be non-negative");
    D2IC_CHECK_ARG(ncand == 0 || cand_elt != nullptr, "cand_elt must not be NULL");
    D2IC_CHECK_ARG(ncand == 0 || n_elems == 0 || (all_set(nodes, elements) && n_nodes > 0),
                   "nodes and elements must not be NULL");
    *n_kept = 0;
    if (ncand == 0) return D2IC_OK;
    if (n_elems == 0) {
        D2IC_HIP(hipMemsetAsync(cand_elt, 0xFF, sizeof(int32_t) * ncand, st));
        return D2IC_OK;
    }
    DevBuf quads, bb;
    D2IC_HIP(quads.alloc(sizeof(Quad) * n_elems, st));
    D2IC_HIP(bb.alloc(4 * sizeof(float), st));
    int tb = 256;
    k_quads<<<ceil_div(n_elems, tb), tb, 0, st>>>(nodes, elements, n_elems, quads.as<Quad>());
    D2IC_LAUNCH_CHECK();
    k_node_bbox<<<1, 256, 0, st>>>(nodes, n_nodes, bb.as<float>());
    D2IC_LAUNCH_CHECK();
    float hb[4];
    D2IC_HIP(hipMemcpyAsync(hb, bb.p, sizeof(hb), hipMemcpyDeviceToHost, st));
    D2IC_HIP(hipStreamSynchronize(st));
    D2IC_CHECK_ARG(std::isfinite(hb[0]) && std::isfinite(hb[1]) && std::isfinite(hb[2]) && std::isfinite(hb[3]),
                   "node coordinates must be finite");
    BinGrid g;
    g.bs = 32.0;
    g.x0 = hb[0];
    g.y0 = hb[1];
    g.nbx = (int)std::floor(((double)hb[2] - g.x0) / g.bs) + 1;
    g.nby = (int)std::floor(((double)hb[3] - g.y0) / g.bs) + 1;
    int64_t nbins = (int64_t)g.nbx * g.nby;
    D2IC_CHECK_ARG(nbins < (int64_t)1 << 28, "mesh extent too large");
    DevBuf cnt, start, fill, list;
    D2IC_HIP(cnt.alloc(sizeof(int32_t) * (nbins + 1), st));
    D2IC_HIP(start.alloc(sizeof(int32_t) * (nbins + 1), st));
    D2IC_HIP(fill.alloc(sizeof(int32_t) * nbins, st));
    D2IC_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (nbins + 1), st));
    D2IC_HIP(hipMemsetAsync(fill.p, 0, sizeof(int32_t) * nbins, st));
    k_bin_count<<<ceil_div(n_elems, tb), tb, 0, st>>>(quads.as<Quad>(), n_elems, g, cnt.as<int32_t>());
    D2IC_LAUNCH_CHECK();
    size_t tmp_bytes = 0;
    D2IC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt.as<int32_t>(), start.as<int32_t>(),
                                              (int)(nbins + 1), st));
    DevBuf tmp;
    D2IC_HIP(tmp.alloc(tmp_bytes, st));
    D2IC_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_bytes, cnt.as<int32_t>(), start.as<int32_t>(),
                                              (int)(nbins + 1), st));
    int32_t total = 0;
    D2IC_HIP(hipMemcpyAsync(&total, start.as<int32_t>() + nbins, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    D2IC_HIP(hipStreamSynchronize(st));
    D2IC_HIP(list.alloc(sizeof(int32_t) * (total + 1), st));
    k_bin_fill<<<ceil_div(n_elems, tb), tb, 0, st>>>(quads.as<Quad>(), n_elems, g, start.as<int32_t>(),
                                                     fill.as<int32_t>(), list.as<int32_t>());
    D2IC_LAUNCH_CHECK();
    k_assign<<<ceil_div(ncand, tb), tb, 0, st>>>(quads.as<Quad>(), g, start.as<int32_t>(), list.as<int32_t>(), H,
                                                 W, i0, j0, (int32_t)nj, ncand, roi, cand_elt);
    D2IC_LAUNCH_CHECK();
    int64_t nblk = ceil_div(ncand, kCompactBlock);
    DevBuf bc;
    D2IC_HIP(bc.alloc(sizeof(int32_t) * nblk, st));
    k_count_kept<<<nblk, kCompactBlock, 0, st>>>(cand_elt, ncand, bc.as<int32_t>());
    D2IC_LAUNCH_CHECK();
    std::vector<int32_t> hc(nblk);
    D2IC_HIP(hipMemcpyAsync(hc.data(), bc.p, sizeof(int32_t) * nblk, hipMemcpyDeviceToHost, st));
    D2IC_HIP(hipStreamSynchronize(st));
    int64_t s = 0;
    for (int64_t k = 0; k < nblk; ++k) s += hc[k];
    *n_kept = s;
    return D2IC_OK;
}

int d2ic_pixel_fill(const int32_t* cand_elt, int32_t H, int32_t W, int32_t i0, int32_t i1, int32_t j0, int32_t j1,
                    const float* nodes, const int32_t* elements, int64_t n_pixels, float* pixel_coords,
                    int32_t* pixel_nodes, float* pixel_shapeN, int32_t* pixel_elements, int32_t* roi_flat,
                    void* stream) {
    hipStream_t st = (hipStream_t)stream;
    D2IC_CHECK_ARG(0 <= i0 && i0 <= i1 && i1 <= H && 0 <= j0 && j0 <= j1 && j1 <= W, "bbox outside image");
    int64_t nj = j1 - j0, ncand = (int64_t)(i1 - i0) * nj;
    D2IC_CHECK_ARG(n_pixels >= 0, "n_pixels must be >= 0");
    if (n_pixels == 0 || ncand == 0) return D2IC_OK;
    D2IC_CHECK_ARG(all_set(cand_elt, nodes, elements, pixel_coords, pixel_nodes, pixel_shapeN, pixel_elements, roi_flat),
                   "arrays must not be NULL");
    int64_t nblk = ceil_div(ncand, kCompactBlock);
    DevBuf bc, bo, src;
    D2IC_HIP(bc.alloc(sizeof(int32_t) * nblk, st));
    D2IC_HIP(bo.alloc(sizeof(int32_t) * nblk, st));
    D2IC_HIP(src.alloc(sizeof(int32_t) * n_pixels, st));
    k_count_kept<<<nblk, kCompactBlock, 0, st>>>(cand_elt, ncand, bc.as<int32_t>());
    D2IC_LAUNCH_CHECK();
    size_t tmp_bytes = 0;
    D2IC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, bc.as<int32_t>(), bo.as<int32_t>(), (int)nblk, st));
    DevBuf tmp;
    D2IC_HIP(tmp.alloc(tmp_bytes, st));
    D2IC_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_bytes, bc.as<int32_t>(), bo.as<int32_t>(), (int)nblk, st));
    // Check the kept count against n_pixels before scattering.
    int32_t last_off = 0, last_cnt = 0;
    D2IC_HIP(hipMemcpyAsync(&last_off, bo.as<int32_t>() + nblk - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    D2IC_HIP(hipMemcpyAsync(&last_cnt, bc.as<int32_t>() + nblk - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    D2IC_HIP(hipStreamSynchronize(st));
    D2IC_CHECK_ARG((int64_t)last_off + last_cnt == n_pixels, "n_pixels (%lld) does not match the kept count (%lld)",
                   (long long)n_pixels, (long long)last_off + last_cnt);
    k_scatter_kept<<<nblk, kCompactBlock, 0, st>>>(cand_elt, ncand, bo.as<int32_t>(), src.as<int32_t>());
    D2IC_LAUNCH_CHECK();
    int tb = 256;
    k_fill_pixels<<<ceil_div(n_pixels, tb), tb, 0, st>>>(src.as<int32_t>(), cand_elt, n_pixels, H, W, i0, j0,
                                                         (int32_t)nj, nodes, elements, pixel_coords, pixel_nodes,
                                                         pixel_shapeN, pixel_elements, roi_flat);
    D2IC_LAUNCH_CHECK();
    return D2IC_OK;
}

int d2ic_node_pixel_degree(const int32_t* pixel_nodes, int64_t n_pixels, int64_t n_nodes, int32_t* degree,
                           int32_t* max_degree, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    D2IC_CHECK_ARG(max_degree != nullptr, "max_degree must not be NULL");
    D2IC_CHECK_ARG(4 * n_pixels < (int64_t)INT32_MAX, "too many pixels");
    D2IC_CHECK_ARG(n_pixels >= 0 && n_nodes >= 0, "sizes must be non-negative");
    D2IC_CHECK_ARG(n_nodes == 0 || degree != nullptr, "degree must not be NULL");
    D2IC_CHECK_ARG(n_nodes == 0 || n_pixels == 0 || pixel_nodes != nullptr, "pixel_nodes must not be NULL");
    *max_degree = 0;
    if (n_nodes == 0) return D2IC_OK;
    D2IC_HIP(hipMemsetAsync(degree, 0, sizeof(int32_t) * n_nodes, st));
    if (n_pixels == 0) return D2IC_OK;
    int tb = 256;
    k_degree<<<ceil_div(4 * n_pixels, tb), tb, 0, st>>>(pixel_nodes, 4 * n_pixels, degree);
    D2IC_LAUNCH_CHECK();
    DevBuf out, tmp;
    D2IC_HIP(out.alloc(sizeof(int32_t), st));
    size_t tmp_bytes = 0;
    D2IC_HIP(hipcub::DeviceReduce::Max(nullptr, tmp_bytes, degree, out.as<int32_t>(), (int)n_nodes, st));
    D2IC_HIP(tmp.alloc(tmp_bytes, st));
    D2IC_HIP(hipcub::DeviceReduce::Max(tmp.p, tmp_bytes, degree, out.as<int32_t>(), (int)n_nodes, st));
    D2IC_HIP(hipMemcpyAsync(max_degree, out.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    D2IC_HIP(hipStreamSynchronize(st));
    return D2IC_OK;
}

int d2ic_node_pixel_table(const int32_t* pixel_nodes, const float* pixel_shapeN, int64_t n_pixels, int64_t n_nodes,
                          int32_t max_degree, int32_t* node_pixel_index, float* node_N_weight, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    D2IC_CHECK_ARG(max_degree >= 0 && n_pixels >= 0 && n_nodes >= 0, "sizes must be non-negative");
    int64_t tab = n_nodes * (int64_t)max_degree;
    if (tab == 0) return D2IC_OK;
    D2IC_CHECK_ARG(all_set(node_pixel_index, node_N_weight), "output tables must not be NULL");
    D2IC_CHECK_ARG(n_pixels == 0 || all_set(pixel_nodes, pixel_shapeN), "pixel arrays must not be NULL");
    D2IC_HIP(hipMemsetAsync(node_pixel_index, 0xFF, sizeof(int32_t) * tab, st));
    D2IC_HIP(hipMemsetAsync(node_N_weight, 0, sizeof(float) * tab, st));
    if (n_pixels == 0) return D2IC_OK;
    int64_t n4 = 4 * n_pixels;
    int tb = 256;
    DevBuf vals, keys_out, vals_out, deg, start, tmp;
    D2IC_HIP(vals.alloc(sizeof(int32_t) * n4, st));
    D2IC_HIP(keys_out.alloc(sizeof(int32_t) * n4, st));
    D2IC_HIP(vals_out.alloc(sizeof(int32_t) * n4, st));
    D2IC_HIP(deg.alloc(sizeof(int32_t) * n_nodes, st));
    D2IC_HIP(start.alloc(sizeof(int32_t) * n_nodes, st));
    k_iota<<<ceil_div(n4, tb), tb, 0, st>>>(vals.as<int32_t>(), n4);
    D2IC_LAUNCH_CHECK();
    // Stable LSD radix sort of the flat (node, pixel*4+a) entries by node keeps
    // each node's entries in ascending flat order = the scan's slot order.
    int bits = bits_for(n_nodes);
    size_t tmp_bytes = 0;
    D2IC_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, pixel_nodes, keys_out.as<int32_t>(),
                                                vals.as<int32_t>(), vals_out.as<int32_t>(), (int)n4, 0, bits, st));
    size_t scan_bytes = 0;
    D2IC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, deg.as<int32_t>(), start.as<int32_t>(),
                                              (int)n_nodes, st));
    D2IC_HIP(tmp.alloc(std::max(tmp_bytes, scan_bytes), st));
    D2IC_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, pixel_nodes, keys_out.as<int32_t>(),
                                                vals.as<int32_t>(), vals_out.as<int32_t>(), (int)n4, 0, bits, st));
    D2IC_HIP(hipMemsetAsync(deg.p, 0, sizeof(int32_t) * n_nodes, st));
    k_degree<<<ceil_div(n4, tb), tb, 0, st>>>(pixel_nodes, n4, deg.as<int32_t>());
    D2IC_LAUNCH_CHECK();
    D2IC_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, scan_bytes, deg.as<int32_t>(), start.as<int32_t>(),
                                              (int)n_nodes, st));
    k_table_fill<<<ceil_div(n4, tb), tb, 0, st>>>(keys_out.as<int32_t>(), vals_out.as<int32_t>(), n4,
                                                  start.as<int32_t>(), pixel_shapeN, max_degree, node_pixel_index,
                                                  node_N_weight);
    D2IC_LAUNCH_CHECK();
    return D2IC_OK;
}

}  // extern "C"
