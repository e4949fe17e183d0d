// Initialisation and post-processing kernels around the solver:
//   * translation ZNCC search at element centres (solver_translation_zncc.py:89-168)
//   * IDW projection of centre shifts to nodes (dic_init_motion.py:64-91)
//   * nodal LSQ Green-Lagrange strain (strain.py:11-114)
//   * reference-domain discrepancy map (discrepancy.py:51-127)
//   * synthetic speckle renderer (workload generator)
#include <hipcub/hipcub.hpp>

#include <vector>

#include "common.hpp"
#include "interp.hpp"

namespace d2ic {

// ---------------------------------------------------------------------------
// ZNCC: one block per centre.  The normalised template and the search patch
// are staged in LDS; kZnccSplit threads score each window (ky, kx);
// the block takes the first maximum (a NaN counts as the maximum, like
// jnp.argmax).  Statistics are accumulated in FP64.
// ---------------------------------------------------------------------------
constexpr int kZnccBlock = 256;
constexpr int kZnccSplit = 3;

__device__ __forceinline__ bool better(float a, int ia, float b, int ib) {
    // is (a, ia) preferred over (b, ib) under "first max, NaN wins"?
    const bool an = isnan(a), bn = isnan(b);
    if (an != bn) return an;
    if (an && bn) return ia < ib;
    if (a != b) return a > b;
    return ia < ib;
}

__global__ __launch_bounds__(kZnccBlock) void k_zncc(const float* __restrict__ centers, const float* __restrict__ ref,
                                                     int Hr, int Wr, const float* __restrict__ def, int Hd, int Wd,
                                                     int win, int search, float* __restrict__ u_out,
                                                     float* __restrict__ score_out) {
    extern __shared__ float lds[];
    const int c = blockIdx.x;
    const int r = win / 2;
    const int P = win + 2 * search;
    const int nwin = 2 * search + 1;
    const int n = win * win;
    const int x = (int)rintf(centers[2 * c]);
    const int y = (int)rintf(centers[2 * c + 1]);
    const bool tpl_ok = (y - r >= 0) && (y + r < Hr) && (x - r >= 0) && (x + r < Wr);
    const int xs = x - search, ys = y - search, xe = x + search, ye = y + search;
    const bool patch_ok = (ys - r >= 0) && (ye + r < Hd) && (xs - r >= 0) && (xe + r < Wd);
    if (!(tpl_ok && patch_ok)) {
        if (threadIdx.x == 0) {
            u_out[2 * c] = 0.f;
            u_out[2 * c + 1] = 0.f;
            score_out[c] = NAN;
        }
        return;
    }
    float* tpl = lds;          // n
    float* patch = lds + n;    // P*P
    __shared__ double red[kZnccBlock];
    __shared__ float s_best[kZnccBlock];
    __shared__ int s_bidx[kZnccBlock];
    const int tr = y - r, tc = x - r;
    const int pr = ys - r, pc = xs - r;
    for (int k = threadIdx.x; k < n; k += blockDim.x) tpl[k] = ref[(size_t)(tr + k / win) * Wr + tc + k % win];
    for (int k = threadIdx.x; k < P * P; k += blockDim.x) patch[k] = def[(size_t)(pr + k / P) * Wd + pc + k % P];
    __syncthreads();
    // template mean / std (population) -> normalised template in place
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) s += tpl[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = kZnccBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    const float tmean = (float)(red[0] / n);
    __syncthreads();
    s = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        float d = tpl[k] - tmean;
        s += (double)(d * d);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = kZnccBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    const float tden = (float)sqrt(red[0] / n) + 1e-12f;
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) tpl[k] = (tpl[k] - tmean) / tden;
    __syncthreads();

    // each candidate window's rows are split over kZnccSplit consecutive
    // threads; the three passes (mean, variance, correlation) combine the
    // split's FP64 partials in a fixed order through LDS
    float best = 0.f;
    int bidx = 0x7fffffff;
    bool have = false;
    const int nc = nwin * nwin;
    constexpr int cap = (kZnccBlock / kZnccSplit) * kZnccSplit;
    const int rows_per = (win + kZnccSplit - 1) / kZnccSplit;
    const int tid = threadIdx.x;
    for (int b0 = 0; b0 < nc * kZnccSplit; b0 += cap) {
        const int item = b0 + tid;
        const bool act = tid < cap && item < nc * kZnccSplit;
        const int wi = act ? item / kZnccSplit : 0, g = act ? item % kZnccSplit : 0;
        const int lead = tid - g;  // thread holding the split's first partial
        const int i0 = min(win, g * rows_per), i1 = min(win, i0 + rows_per);
        const int ky = wi / nwin, kx = wi % nwin;
        const float* base = patch + ky * P + kx;
        double sw = 0.0;
        if (act)
            for (int i = i0; i < i1; ++i)
#pragma unroll 4
                for (int j = 0; j < win; ++j) sw += base[i * P + j];
        red[tid] = sw;
        __syncthreads();
        double tot = 0.0;
        if (act)
            for (int q = 0; q < kZnccSplit; ++q) tot += red[lead + q];
        const float m = (float)(tot / n);
        __syncthreads();
        double sv = 0.0;
        if (act)
            for (int i = i0; i < i1; ++i)
#pragma unroll 4
                for (int j = 0; j < win; ++j) {
                    float d = base[i * P + j] - m;
                    sv += (double)(d * d);
                }
        red[tid] = sv;
        __syncthreads();
        tot = 0.0;
        if (act)
            for (int q = 0; q < kZnccSplit; ++q) tot += red[lead + q];
        const float den = (float)sqrt(tot / n) + 1e-12f;
        __syncthreads();
        double acc = 0.0;
        if (act)
            for (int i = i0; i < i1; ++i)
#pragma unroll 4
                for (int j = 0; j < win; ++j) acc += (double)(((base[i * P + j] - m) / den) * tpl[i * win + j]);
        red[tid] = acc;
        __syncthreads();
        if (act && g == 0) {
            tot = 0.0;
            for (int q = 0; q < kZnccSplit; ++q) tot += red[lead + q];
            const float sc = (float)(tot / n);
            if (!have || better(sc, wi, best, bidx)) {
                best = sc;
                bidx = wi;
                have = true;
            }
        }
        __syncthreads();
    }
    s_best[threadIdx.x] = best;
    s_bidx[threadIdx.x] = have ? bidx : 0x7fffffff;
    __syncthreads();
    for (int off = kZnccBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            int ob = s_bidx[threadIdx.x + off];
            if (ob != 0x7fffffff &&
                (s_bidx[threadIdx.x] == 0x7fffffff ||
                 better(s_best[threadIdx.x + off], ob, s_best[threadIdx.x], s_bidx[threadIdx.x]))) {
                s_best[threadIdx.x] = s_best[threadIdx.x + off];
                s_bidx[threadIdx.x] = ob;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int bi = s_bidx[0];
        const int ky = bi / nwin, kx = bi % nwin;
        u_out[2 * c] = (float)((xs + kx) - x);
        u_out[2 * c + 1] = (float)((ys + ky) - y);
        score_out[c] = s_best[0];
    }
}

// ---------------------------------------------------------------------------
// IDW (power 2) with exact-hit copy.  A block serves 64 nodes (one per
// lane); its kIdwWaves waves take interleaved 64-centre tiles, each staged in
// the wave's own LDS slot, and the waves' FP64 partial sums are added in wave
// order.  The copied centre is the first one within 1e-6 (smallest index):
// each wave stops at its own first hit and the smallest hit index wins.
// ---------------------------------------------------------------------------
constexpr int kIdwWaves = 8;

__global__ __launch_bounds__(64 * kIdwWaves) void k_idw(const float* __restrict__ nodes, int64_t nn,
                                                        const float* __restrict__ centers,
                                                        const float* __restrict__ disp, int64_t nc,
                                                        float* __restrict__ out) {
    __shared__ float4 sc[kIdwWaves][64];
    __shared__ double part[3][kIdwWaves][64];
    __shared__ int64_t hit[kIdwWaves][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t n = (int64_t)blockIdx.x * 64 + lane;
    const bool valid = n < nn;
    const float nx = valid ? nodes[2 * n] : 0.f, ny = valid ? nodes[2 * n + 1] : 0.f;
    double ws = 0.0, ax = 0.0, ay = 0.0;
    int64_t close = -1;
    for (int64_t c0 = (int64_t)w * 64; c0 < nc; c0 += 64 * kIdwWaves) {
        // LDS operations of one wave complete in order: the previous tile's
        // reads are done before these writes, and these writes before the reads
        const int64_t ci = c0 + lane;
        __builtin_amdgcn_wave_barrier();
        if (ci < nc) sc[w][lane] = make_float4(centers[2 * ci], centers[2 * ci + 1], disp[2 * ci], disp[2 * ci + 1]);
        __builtin_amdgcn_wave_barrier();
        const int m = (int)min<int64_t>(64, nc - c0);
        if (valid && close < 0) {
            for (int k = 0; k < m; ++k) {
                const float4 q = sc[w][k];
                const float dx = q.x - nx, dy = q.y - ny;
                const float dist = sqrtf(dx * dx + dy * dy);
                if (dist < 1e-6f) {
                    close = c0 + k;
                    break;
                }
                const float dd = dist + 1e-12f;
                const float wgt = 1.0f / (dd * dd);
                ws += wgt;
                ax += (double)wgt * q.z;
                ay += (double)wgt * q.w;
            }
        }
    }
    part[0][w][lane] = ws;
    part[1][w][lane] = ax;
    part[2][w][lane] = ay;
    hit[w][lane] = close;
    __syncthreads();
    if (w != 0 || !valid) return;
    ws = ax = ay = 0.0;
    close = -1;
    for (int q = 0; q < kIdwWaves; ++q) {
        ws += part[0][q][lane];
        ax += part[1][q][lane];
        ay += part[2][q][lane];
        const int64_t h = hit[q][lane];
        if (h >= 0 && (close < 0 || h < close)) close = h;
    }
    if (close >= 0) {
        out[2 * n] = disp[2 * close];
        out[2 * n + 1] = disp[2 * close + 1];
    } else {
        out[2 * n] = (float)(ax / ws);
        out[2 * n + 1] = (float)(ay / ws);
    }
}

// ---------------------------------------------------------------------------
// Nodal weighted LSQ gradient and Green-Lagrange strain.
// ---------------------------------------------------------------------------
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

__global__ void k_strain(const float* __restrict__ u, const float* __restrict__ X, int64_t nn,
                         const int32_t* __restrict__ nbi, const int32_t* __restrict__ nbd, int md, float L, float eps,
                         float* __restrict__ Fo, float* __restrict__ Eo, float* __restrict__ V) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    float a00 = 0.f, a01 = 0.f, a10 = 0.f, a11 = 0.f, b0x = 0.f, b0y = 0.f, b1x = 0.f, b1y = 0.f;
    const int deg = nbd[i];
    const float xi = X[2 * i], yi = X[2 * i + 1], uix = u[2 * i], uiy = u[2 * i + 1];
    for (int k = 0; k < md; ++k) {
        const bool valid = k < deg;
        int j = valid ? nbi[i * md + k] : 0;
        if (j < 0 || j >= nn) j = 0;
        const float dX = X[2 * j] - xi, dY = X[2 * j + 1] - yi;
        const float dux = u[2 * j] - uix, duy = u[2 * j + 1] - uiy;
        float w = valid ? 1.f : 0.f;
        if (L > 0.f) {
            const float rr = sqrtf(dX * dX + dY * dY);
            const float z = rr / (L + 1e-12f);
            w = w * expf(-(z * z));
        }
        const float wx = dX * w, wy = dY * w;
        a00 += wx * dX;
        a01 += wx * dY;
        a10 += wy * dX;
        a11 += wy * dY;
        b0x += wx * dux;
        b0y += wy * dux;
        b1x += wx * duy;
        b1y += wy * duy;
    }
    a00 += eps;
    a11 += eps;
    float gxx, gxy, gyx, gyy;
    solve2(a00, a01, a10, a11, b0x, b0y, gxx, gxy);
    solve2(a00, a01, a10, a11, b1x, b1y, gyx, gyy);
    const float F00 = 1.f + gxx, F01 = gxy, F10 = gyx, F11 = 1.f + gyy;
    const float C00 = F00 * F00 + F10 * F10, C01 = F00 * F01 + F10 * F11;
    const float C10 = F01 * F00 + F11 * F10, C11 = F01 * F01 + F11 * F11;
    const float E00 = 0.5f * (C00 - 1.f), E01 = 0.5f * C01, E10 = 0.5f * C10, E11 = 0.5f * (C11 - 1.f);
    if (Fo) {
        Fo[4 * i] = F00; Fo[4 * i + 1] = F01; Fo[4 * i + 2] = F10; Fo[4 * i + 3] = F11;
    }
    if (Eo) {
        Eo[4 * i] = E00; Eo[4 * i + 1] = E01; Eo[4 * i + 2] = E10; Eo[4 * i + 3] = E11;
    }
    V[3 * i] = E00;
    V[3 * i + 1] = E11;
    V[3 * i + 2] = 0.5f * (E01 + E10);
}

// ---------------------------------------------------------------------------
// Discrepancy map (discrepancy.py:86-127).  bilinear_sample_jax (:51-83) at x
// and x + u(x) with NaN outside [0, H-1) x [0, W-1); the residual is scattered
// like jnp.bincount(roi_flat, weights=r * valid) / max(bincount(valid), 1), NaN
// where no valid pixel landed.  Every operation is a separately rounded FP32
// operation in the reference's order (no contraction), so a pixel's residual
// is the reference's FP32 value.  A NaN residual adds NaN to its image pixel
// (the reference's NaN * 0 weight); per-block FP64 partials give the RMS.
// (hipcc contracts a * b + c into an FMA by default, also inside the
// __fmul_rn / __fadd_rn helpers, so these bodies switch contraction off and
// use plain operators.)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float bilin_nan(const float* img, int H, int W, float x, float y) {
#pragma clang fp contract(off)
    const float xc = x - 0.5f, yc = y - 0.5f;
    const int j0 = (int)floorf(xc), i0 = (int)floorf(yc);
    const float fx = xc - (float)j0, fy = yc - (float)i0;
    if (!(i0 >= 0 && i0 < H - 1 && j0 >= 0 && j0 < W - 1)) return NAN;
    const float v00 = img[(size_t)i0 * W + j0], v01 = img[(size_t)i0 * W + j0 + 1];
    const float v10 = img[(size_t)(i0 + 1) * W + j0], v11 = img[(size_t)(i0 + 1) * W + j0 + 1];
    const float gx = 1.0f - fx, gy = 1.0f - fy;
    float v = v00 * gx * gy;
    v = v + v01 * fx * gy;
    v = v + v10 * gx * fy;
    v = v + v11 * fx * fy;
    return v;
}

__global__ void k_nan_fill(float* __restrict__ m, int64_t n) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) m[t] = NAN;
}

// zero sums and counts (a fill kernel rather than a memset node, so that a
// caller may capture the launch sequence into a graph)
__global__ void k_disc_clear(float* __restrict__ sum, int32_t* __restrict__ cnt, int64_t n) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        sum[t] = 0.f;
        cnt[t] = 0;
    }
}

// FP32 add by compare-and-swap: IEEE round-to-nearest with denormals and NaN
// propagated, as a sequential bincount would (one CAS per pixel when the
// image pixels are distinct, the case of every asset builder).
__device__ __forceinline__ void cas_add(float* p, float v) {
    unsigned* q = reinterpret_cast<unsigned*>(p);
    unsigned old = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (true) {
        const unsigned nw = __float_as_uint(__uint_as_float(old) + v);
        const unsigned seen = atomicCAS(q, old, nw);
        if (seen == old) break;
        old = seen;
    }
}

__global__ __launch_bounds__(256) void k_discrepancy(const float* __restrict__ ref, const float* __restrict__ def,
                                                     int H, int W, const float* __restrict__ u,
                                                     const float* __restrict__ xy, const int32_t* __restrict__ pn,
                                                     const float* __restrict__ N, const int32_t* __restrict__ rf,
                                                     int64_t np, float* __restrict__ sum, int32_t* __restrict__ cnt,
                                                     double* __restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double s0[256], s1[256], s2[256];
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0, b = 0.0, c = 0.0;
    if (p < np) {
        // pixel displacement: sum_a N_a u_a in node order (np.sum over axis 1)
        float ux = 0.f, uy = 0.f;
        for (int k = 0; k < 4; ++k) {
            const int n = pn[4 * p + k];
            const float tx = N[4 * p + k] * u[2 * n], ty = N[4 * p + k] * u[2 * n + 1];
            ux = k == 0 ? tx : ux + tx;
            uy = k == 0 ? ty : uy + ty;
        }
        const float x = xy[2 * p], y = xy[2 * p + 1];
        const float ir = bilin_nan(ref, H, W, x, y);
        const float id = bilin_nan(def, H, W, x + ux, y + uy);
        const float r = id - ir;
        const bool valid = isfinite(r);
        const int64_t idx = rf[p];
        if (idx >= 0 && idx < (int64_t)H * W) {
            cas_add(sum + idx, valid ? r : NAN);
            if (valid) atomicAdd(cnt + idx, 1);
        }
        if (valid) {
            a = (double)r * (double)r;
            b = 1.0;
        } else {
            c = 1.0;
        }
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = b;
    s2[threadIdx.x] = c;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            s0[threadIdx.x] += s0[threadIdx.x + off];
            s1[threadIdx.x] += s1[threadIdx.x + off];
            s2[threadIdx.x] += s2[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = s0[0];
        part[3 * blockIdx.x + 1] = s1[0];
        part[3 * blockIdx.x + 2] = s2[0];
    }
}

// map = sum / count where a valid pixel landed, NaN elsewhere
__global__ void k_discrepancy_finish(float* __restrict__ map, const int32_t* __restrict__ cnt, int64_t n) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int c = cnt[t];
    map[t] = c > 0 ? __fdiv_rn(map[t], (float)c) : NAN;
}

// ---------------------------------------------------------------------------
// Speckle renderer: spots sorted by 16-px bin (stable), each pixel sums the
// spots of the 3x3 neighbouring bins in sorted order (deterministic).
// ---------------------------------------------------------------------------
constexpr int kSpotBin = 16;

__global__ void k_spot_keys(const double* __restrict__ spots, int64_t n, int nbx, int nby, int32_t* __restrict__ keys,
                            int32_t* __restrict__ vals) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int bx = (int)floor(spots[4 * k] / kSpotBin), by = (int)floor(spots[4 * k + 1] / kSpotBin);
    bx = clampi(bx, 0, nbx - 1);
    by = clampi(by, 0, nby - 1);
    keys[k] = by * nbx + bx;
    vals[k] = (int32_t)k;
}

__global__ void k_bin_bounds(const int32_t* __restrict__ keys, int64_t n, int nbins, int32_t* __restrict__ start) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > n) return;
    int cur = k < n ? keys[k] : nbins;
    int prev = k > 0 ? keys[k - 1] : -1;
    for (int b = prev + 1; b <= cur; ++b) start[b] = (int32_t)k;
}

__global__ void k_render(const double* __restrict__ spots, const int32_t* __restrict__ order,
                         const int32_t* __restrict__ start, int nbx, int nby, int H, int W, double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)H * W) return;
    const int i = (int)(t / W), j = (int)(t % W);
    const double px = j + 0.5, py = i + 0.5;
    const int bx = clampi(j / kSpotBin, 0, nbx - 1), by = clampi(i / kSpotBin, 0, nby - 1);
    double acc = 0.0;
    for (int yy = max(0, by - 1); yy <= min(nby - 1, by + 1); ++yy)
        for (int xx = max(0, bx - 1); xx <= min(nbx - 1, bx + 1); ++xx) {
            const int b = yy * nbx + xx;
            for (int k = start[b]; k < start[b + 1]; ++k) {
                const int s = order[k];
                const double dx = px - spots[4 * s], dy = py - spots[4 * s + 1];
                const double sg = spots[4 * s + 2];
                const double r2 = dx * dx + dy * dy;
                if (r2 <= 16.0 * sg * sg) acc += spots[4 * s + 3] * exp(-r2 / (2.0 * sg * sg));
            }
        }
    out[t] = acc;
}

}  // namespace d2ic

using namespace d2ic;

namespace {
struct Tmp {
    void* p = nullptr;
    hipStream_t s = nullptr;
    ~Tmp() {
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
}  // namespace

extern "C" {

int d2ic_zncc_match(const float* centers, int64_t n_centers, const float* ref, int32_t Hr, int32_t Wr,
                    const float* def, int32_t Hd, int32_t Wd, int32_t win, int32_t search, float* u_out,
                    float* score_out, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    D2IC_CHECK_ARG(win > 0 && win % 2 == 1, "win must be odd and positive");
    D2IC_CHECK_ARG(search >= 0, "search must be >= 0");
    D2IC_CHECK_ARG(Hr > 0 && Wr > 0 && Hd > 0 && Wd > 0, "images must be non-empty");
    D2IC_CHECK_ARG(n_centers >= 0, "n_centers must be >= 0");
    if (n_centers == 0) return D2IC_OK;
    D2IC_CHECK_ARG(all_set(centers, ref, def, u_out, score_out), "arrays must not be NULL");
    const int P = win + 2 * search;
    size_t lds = sizeof(float) * ((size_t)win * win + (size_t)P * P);
    D2IC_CHECK_ARG(lds <= 60 * 1024, "win + 2*search too large for the LDS-staged search (%d)", P);
    k_zncc<<<(unsigned)n_centers, kZnccBlock, lds, st>>>(centers, ref, Hr, Wr, def, Hd, Wd, win, search, u_out,
                                                         score_out);
    D2IC_LAUNCH_CHECK();
    return D2IC_OK;
}

int d2ic_idw(const float* nodes, int64_t n_nodes, const float* centers, const float* disp, int64_t n_centers,
             float* u_out, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    D2IC_CHECK_ARG(n_nodes >= 0 && n_centers >= 0, "sizes must be non-negative");
    if (n_nodes == 0) return D2IC_OK;
    D2IC_CHECK_ARG(all_set(nodes, u_out), "nodes and u_out must not be NULL");
    D2IC_CHECK_ARG(n_centers == 0 || all_set(centers, disp), "centers and disp must not be NULL");
    if (n_centers == 0) {
        D2IC_HIP(hipMemsetAsync(u_out, 0, sizeof(float) * 2 * n_nodes, st));
        return D2IC_OK;
    }
    k_idw<<<ceil_div(n_nodes, 64), 64 * kIdwWaves, 0, st>>>(nodes, n_nodes, centers, disp, n_centers, u_out);
    D2IC_LAUNCH_CHECK();
    return D2IC_OK;
}

int d2ic_strain(const float* u, const float* nodes, int64_t n_nodes, const int32_t* nb_index,
                const int32_t* nb_degree, int32_t max_degree, float gauge_length, float eps, float* F_out,
                float* E_out, float* voigt, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    D2IC_CHECK_ARG(voigt != nullptr, "voigt must not be NULL");
    D2IC_CHECK_ARG(n_nodes >= 0 && max_degree >= 0, "sizes must be non-negative");
    if (n_nodes == 0) return D2IC_OK;
    D2IC_CHECK_ARG(all_set(u, nodes, nb_degree, F_out, E_out), "arrays must not be NULL");
    D2IC_CHECK_ARG(max_degree == 0 || nb_index != nullptr, "nb_index must not be NULL");
    k_strain<<<ceil_div(n_nodes, 256), 256, 0, st>>>(u, nodes, n_nodes, nb_index, nb_degree, max_degree,
                                                     gauge_length, eps, F_out, E_out, voigt);
    D2IC_LAUNCH_CHECK();
    return D2IC_OK;
}

int d2ic_discrepancy_map(const float* ref_image, const float* def_image, int32_t H, int32_t W, const float* u,
                         const float* pixel_coords, const int32_t* pixel_nodes, const float* pixel_shapeN,
                         const int32_t* roi_flat, int64_t n_pixels, float* map_out, float* rms, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    D2IC_CHECK_ARG(H > 0 && W > 0, "image shape must be positive");
    D2IC_CHECK_ARG(n_pixels >= 0, "n_pixels must be >= 0");
    D2IC_CHECK_ARG(map_out != nullptr, "map_out must not be NULL");
    D2IC_CHECK_ARG(n_pixels == 0 || all_set(ref_image, def_image, u, pixel_coords, pixel_nodes, pixel_shapeN, roi_flat),
                   "arrays must not be NULL");
    int64_t hw = (int64_t)H * W;
    if (n_pixels == 0) {
        k_nan_fill<<<ceil_div(hw, 256), 256, 0, st>>>(map_out, hw);
        D2IC_LAUNCH_CHECK();
        if (rms) *rms = 0.f;
        return D2IC_OK;
    }
    int64_t nb = ceil_div(n_pixels, 256);
    Tmp part, cnt;
    D2IC_HIP(part.alloc(sizeof(double) * 3 * nb, st));
    D2IC_HIP(cnt.alloc(sizeof(int32_t) * hw, st));
    k_disc_clear<<<ceil_div(hw, 256), 256, 0, st>>>(map_out, cnt.as<int32_t>(), hw);
    D2IC_LAUNCH_CHECK();
    k_discrepancy<<<nb, 256, 0, st>>>(ref_image, def_image, H, W, u, pixel_coords, pixel_nodes, pixel_shapeN,
                                      roi_flat, n_pixels, map_out, cnt.as<int32_t>(), part.as<double>());
    D2IC_LAUNCH_CHECK();
    k_discrepancy_finish<<<ceil_div(hw, 256), 256, 0, st>>>(map_out, cnt.as<int32_t>(), hw);
    D2IC_LAUNCH_CHECK();
    if (rms) {
        std::vector<double> h(3 * nb);
        D2IC_HIP(hipMemcpyAsync(h.data(), part.p, sizeof(double) * 3 * nb, hipMemcpyDeviceToHost, st));
        D2IC_HIP(hipStreamSynchronize(st));
        double a = 0.0, b = 0.0, c = 0.0;
        for (int64_t k = 0; k < nb; ++k) {
            a += h[3 * k];
            b += h[3 * k + 1];
            c += h[3 * k + 2];
        }
        // the reference sums (r * valid)^2 with NaN * 0 = NaN for an invalid
        // pixel, so one sample outside the domain makes the RMS NaN
        *rms = c > 0.0 ? NAN : (float)std::sqrt(a / std::max(b, 1.0));
    }
    return D2IC_OK;
}

int d2ic_render_speckle(const double* spots, int64_t n_spots, int32_t H, int32_t W, double* out, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    D2IC_CHECK_ARG(H > 0 && W > 0, "image shape must be positive");
    D2IC_CHECK_ARG(out != nullptr, "out must not be NULL");
    D2IC_CHECK_ARG(n_spots >= 0 && (n_spots == 0 || spots != nullptr), "spots must not be NULL");
    const int nbx = (W + kSpotBin - 1) / kSpotBin, nby = (H + kSpotBin - 1) / kSpotBin, nbins = nbx * nby;
    Tmp keys, vals, keys_s, order, start, tmp;
    const int64_t n = n_spots;
    D2IC_HIP(keys.alloc(sizeof(int32_t) * (n + 1), st));
    D2IC_HIP(vals.alloc(sizeof(int32_t) * (n + 1), st));
    D2IC_HIP(keys_s.alloc(sizeof(int32_t) * (n + 1), st));
    D2IC_HIP(order.alloc(sizeof(int32_t) * (n + 1), st));
    D2IC_HIP(start.alloc(sizeof(int32_t) * (nbins + 1), st));
    if (n > 0) {
        k_spot_keys<<<ceil_div(n, 256), 256, 0, st>>>(spots, n, nbx, nby, keys.as<int32_t>(), vals.as<int32_t>());
        D2IC_LAUNCH_CHECK();
        size_t b = 0;
        int bits = 1;
        while ((1 << bits) < nbins + 1) ++bits;
        D2IC_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b, keys.as<int32_t>(), keys_s.as<int32_t>(),
                                                    vals.as<int32_t>(), order.as<int32_t>(), (int)n, 0, bits, st));
        D2IC_HIP(tmp.alloc(b, st));
        D2IC_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, b, keys.as<int32_t>(), keys_s.as<int32_t>(),
                                                    vals.as<int32_t>(), order.as<int32_t>(), (int)n, 0, bits, st));
    }
    k_bin_bounds<<<ceil_div(n + 1, 256), 256, 0, st>>>(keys_s.as<int32_t>(), n, nbins, start.as<int32_t>());
    D2IC_LAUNCH_CHECK();
    k_render<<<ceil_div((int64_t)H * W, 256), 256, 0, st>>>(spots, order.as<int32_t>(), start.as<int32_t>(), nbx, nby,
                                                             H, W, out);
    D2IC_LAUNCH_CHECK();
    return D2IC_OK;
}

}  // extern "C"
