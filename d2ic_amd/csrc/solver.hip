// Global FE-DIC solver: NLCG (Polak-Ribiere+, Armijo backtracking) on
//   J(u) = 1/2 mean_p (g(x_p + u(x_p)) - f(x_p))^2 + alpha_reg * E_spring(u)
// (reference: d2ic/solver_global_cg.py:205-469), as ONE persistent cooperative
// kernel that runs every iteration of every problem of a batch.
//
// Layout (built once per PixelAssets by d2ic_cg_plan_create):
//   * pixels grouped by element and cut into tiles of <= 64 pixels; per pixel
//     structure-of-arrays x, y, N0..N3 (float4) and the pre-sampled f(x_p);
//     per tile its pixel range and its 4 node ids;
//   * a node -> (tile, local node) gather table, so the image part of the
//     gradient is assembled deterministically (no float atomics);
//   * the spring stencil with the combined weight w_ij * (c_i + c_j) / 2.
//
// Each iteration is two grid-wide steps separated by a grid barrier:
//   A-step: every tile evaluates the residual at the candidate
//           u + alpha * d (d formed on the fly from g, d_prev and beta) and
//           writes its partial gradient (8 floats) and its sum of r^2; node
//           chunks materialise d, and accumulate g.d and the spring energy at
//           the candidate.
//   B-step: the Armijo decision is taken (redundantly and identically in
//           every block); if the trial is accepted, node chunks assemble the
//           gradient at the new point through the gather table and reduce
//           |g|^2 and g.(g - g_prev) for the next beta and the stop test.
// The accepted trial's A-step is also the next iteration's value+gradient
// evaluation, so one pixel pass per iteration is enough when alpha = 1 is
// accepted (the reference's usual case), without changing the iterates.
//
// Reductions are deterministic and independent of the batch composition:
// per-tile sums of r^2 are converted exactly to a two-word fixed-point integer
// (resolution 2^-40) and added with integer atomics; node-chunk partials are
// reduced in fixed order by the last chunk to finish.  Dot products and the
// objective are accumulated in FP64 over FP32 arithmetic.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "interp.hpp"

namespace d2ic {

constexpr int kTile = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kMaxProblems = 256;
constexpr double kFixScale = 1099511627776.0;  // 2^40

enum Mode : int { MODE_INIT = 0, MODE_PASS = 1, MODE_NODE = 2, MODE_DONE = 3 };

enum Flag : int {
    F_DOA = 1,
    F_DOB = 2,
    F_RESTART = 4,
    F_CUR = 16,
    F_DCUR = 32,
};

struct ProbDesc {
    const int4* tile_hdr;    // {pixel start, count, group, 0}
    const int4* tile_nodes;  // 4 node ids
    const float* px;
    const float* py;
    const float4* N4;
    const float* F;
    const int32_t* node_ptr;
    const int32_t* node_ent;
    const int32_t* nb_idx;
    const float* nb_wp;
    float* U[2];
    float* G[2];
    float* D[2];
    float* TP;
    double* NCP;
    const float* img;
    float* hist;
    int32_t* n_iters;
    int32_t* pass_count;
    int H, W, interp, nb_maxdeg;
    int ntiles, nnodes, nchunks, max_iter;
    float alpha_reg, tol;
    double invNp;
};

struct ProbState {
    int mode, k, trial, restart, final_eval, cur, dcur, passes;
    float alpha, beta, gnorm, Jtrial;
    double J, gg;
};

struct Accum {
    unsigned long long hi[kMaxProblems];
    unsigned long long lo[kMaxProblems];
    unsigned cnt[kMaxProblems];
    int bad[kMaxProblems];
    double fin[kMaxProblems][2];
};

struct Barrier {
    unsigned count;
    unsigned gen;
    int err;
    int pad;
};

// ---------------------------------------------------------------------------
// Grid barrier with a bounded wait: on timeout (or when another block timed
// out) it returns false and the caller leaves its loop, so every wave exits.
// ---------------------------------------------------------------------------
__device__ bool grid_sync(Barrier* b, unsigned nblocks, uint64_t timeout_ticks) {
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        unsigned g = __hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        unsigned arrived = atomicAdd(&b->count, 1u);
        if (arrived == nblocks - 1) {
            atomicExch(&b->count, 0u);
            __threadfence();
            atomicAdd(&b->gen, 1u);
        } else {
            const uint64_t t0 = wall_clock64();
            while (__hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                if (__hip_atomic_load(&b->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    ok = 0;
                    break;
                }
                if (wall_clock64() - t0 > timeout_ticks) {
                    atomicExch(&b->err, 1);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __threadfence();
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

__device__ __forceinline__ float2 ld2(const float* p, int n) { return reinterpret_cast<const float2*>(p)[n]; }
__device__ __forceinline__ void st2(float* p, int n, float2 v) { reinterpret_cast<float2*>(p)[n] = v; }

// Search direction d = -g (restart) or -g + beta * d_prev, single rounding per
// component so every place that forms it gets the same bits.
__device__ __forceinline__ float2 direction(float2 g, float2 dp, float beta, bool restart) {
    if (restart) return make_float2(-g.x, -g.y);
    return make_float2(__fmaf_rn(beta, dp.x, -g.x), __fmaf_rn(beta, dp.y, -g.y));
}

__device__ __forceinline__ float2 step_to(float2 u, float2 d, float alpha) {
    return make_float2(__fmaf_rn(alpha, d.x, u.x), __fmaf_rn(alpha, d.y, u.y));
}

// Exact split of a non-negative double (resolution 2^-40) into integer and
// fractional fixed-point words.
__device__ __forceinline__ void fix_split(double v, unsigned long long& hi, unsigned long long& lo) {
    double ip = floor(v);
    double fr = (v - ip) * kFixScale;
    hi += (unsigned long long)ip;
    lo += (unsigned long long)fr;
}

__device__ __forceinline__ int find_prob(const int* pref, int P, int item) {
    int lo = 0, hi = P - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (pref[mid] <= item)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Last node chunk of a problem to finish reduces the chunk partials in fixed order.
__device__ void chunk_done(double v0, double v1, double* NCP, int c, int nchunks, Accum* acc, int p, int lane) {
    if (lane == 0) {
        NCP[2 * c] = v0;
        NCP[2 * c + 1] = v1;
    }
    __threadfence();
    unsigned old = 0;
    if (lane == 0) old = atomicAdd(&acc->cnt[p], 1u);
    old = __shfl(old, 0, kWave);
    if (old == (unsigned)(nchunks - 1)) {
        __threadfence();
        double s0 = 0.0, s1 = 0.0;
        for (int k = lane; k < nchunks; k += kWave) {
            s0 += __hip_atomic_load(&NCP[2 * k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s1 += __hip_atomic_load(&NCP[2 * k + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s0 = wave_sum(s0);
        s1 = wave_sum(s1);
        if (lane == 0) {
            acc->fin[p][0] = s0;
            acc->fin[p][1] = s1;
        }
    }
}

// A-step, one tile of <= 64 pixels per wave.
__device__ void tile_pass(const ProbDesc* __restrict__ D, int t, int lane, int flags, float alpha, float beta,
                          unsigned long long& hi, unsigned long long& lo, int& bad) {
    const int cur = (flags & F_CUR) ? 1 : 0, dcur = (flags & F_DCUR) ? 1 : 0;
    const bool restart = (flags & F_RESTART) != 0;
    const int4 hdr = D->tile_hdr[t];
    const int4 nd = D->tile_nodes[t];
    const float* U = D->U[cur];
    const float* G = D->G[cur];
    const float* Dp = D->D[dcur];
    const int nid[4] = {nd.x, nd.y, nd.z, nd.w};
    float2 uc[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int n = nid[a];
        const float2 d = direction(ld2(G, n), ld2(Dp, n), beta, restart);
        uc[a] = step_to(ld2(U, n), d, alpha);
    }
    const bool valid = lane < hdr.y;
    float r = 0.f, gx = 0.f, gy = 0.f;
    float4 N = make_float4(0.f, 0.f, 0.f, 0.f);
    if (valid) {
        const int i = hdr.x + lane;
        const float x = D->px[i], y = D->py[i];
        N = D->N4[i];
        const float f = D->F[i];
        const float ux = N.x * uc[0].x + N.y * uc[1].x + N.z * uc[2].x + N.w * uc[3].x;
        const float uy = N.x * uc[0].y + N.y * uc[1].y + N.z * uc[2].y + N.w * uc[3].y;
        float gv;
        sample(D->interp, D->img, D->H, D->W, x + ux, y + uy, gv, gx, gy);
        r = gv - f;
    }
    const double r2 = wave_sum((double)r * (double)r);
    if (lane == 0) {
        if (!(r2 <= 1e300))
            bad = 1;
        else
            fix_split(r2, hi, lo);
    }
    const float qx = r * gx, qy = r * gy;
    float s0 = wave_sum(qx * N.x), s1 = wave_sum(qy * N.x);
    float s2 = wave_sum(qx * N.y), s3 = wave_sum(qy * N.y);
    float s4 = wave_sum(qx * N.z), s5 = wave_sum(qy * N.z);
    float s6 = wave_sum(qx * N.w), s7 = wave_sum(qy * N.w);
    if (lane == 0) {
        float4* o = reinterpret_cast<float4*>(D->TP + 8 * (size_t)t);
        o[0] = make_float4(s0, s1, s2, s3);
        o[1] = make_float4(s4, s5, s6, s7);
    }
}

// A-step, one chunk of 64 nodes per wave: d, g.d and the spring energy at the candidate.
__device__ void node_pass_a(const ProbDesc* __restrict__ D, int c, int lane, int flags, float alpha, float beta,
                            Accum* acc, int p) {
    const int cur = (flags & F_CUR) ? 1 : 0, dcur = (flags & F_DCUR) ? 1 : 0;
    const bool restart = (flags & F_RESTART) != 0;
    const int n = c * kWave + lane;
    double gd = 0.0, E = 0.0;
    if (n < D->nnodes) {
        const float* U = D->U[cur];
        const float* G = D->G[cur];
        const float* Dp = D->D[dcur];
        const float2 g = ld2(G, n);
        const float2 d = direction(g, ld2(Dp, n), beta, restart);
        st2(D->D[dcur ^ 1], n, d);
        gd = (double)g.x * (double)d.x + (double)g.y * (double)d.y;
        const float2 ucn = step_to(ld2(U, n), d, alpha);
        const int md = D->nb_maxdeg;
        for (int k = 0; k < md; ++k) {
            const int j = D->nb_idx[(size_t)n * md + k];
            if (j < 0) break;
            const float w = D->nb_wp[(size_t)n * md + k];
            const float2 dj = direction(ld2(G, j), ld2(Dp, j), beta, restart);
            const float2 ucj = step_to(ld2(U, j), dj, alpha);
            const float dx = ucn.x - ucj.x, dy = ucn.y - ucj.y;
            E += (double)(w * (dx * dx + dy * dy));
        }
    }
    gd = wave_sum(gd);
    E = wave_sum(E);
    chunk_done(gd, E, D->NCP, c, D->nchunks, acc, p, lane);
}

// B-step, one chunk of 64 nodes per wave: new point, gradient gather, dots.
__device__ void node_pass_b(const ProbDesc* __restrict__ D, int c, int lane, int flags, float alpha, Accum* acc,
                            int p) {
    const int cur = (flags & F_CUR) ? 1 : 0, dcur = (flags & F_DCUR) ? 1 : 0;
    const int n = c * kWave + lane;
    double gg = 0.0, gnum = 0.0;
    if (n < D->nnodes) {
        const float* U = D->U[cur];
        const float* Dn = D->D[dcur ^ 1];
        const float2 un = step_to(ld2(U, n), ld2(Dn, n), alpha);
        st2(D->U[cur ^ 1], n, un);
        double sx = 0.0, sy = 0.0;
        const int e0 = D->node_ptr[n], e1 = D->node_ptr[n + 1];
        for (int e = e0; e < e1; ++e) {
            const float2 q = ld2(D->TP, D->node_ent[e]);
            sx += (double)q.x;
            sy += (double)q.y;
        }
        float rx = 0.f, ry = 0.f;
        const int md = D->nb_maxdeg;
        for (int k = 0; k < md; ++k) {
            const int j = D->nb_idx[(size_t)n * md + k];
            if (j < 0) break;
            const float w = D->nb_wp[(size_t)n * md + k];
            const float2 uj = step_to(ld2(U, j), ld2(Dn, j), alpha);
            rx += w * (un.x - uj.x);
            ry += w * (un.y - uj.y);
        }
        const float gx = (float)(sx * D->invNp) + D->alpha_reg * rx;
        const float gy = (float)(sy * D->invNp) + D->alpha_reg * ry;
        const float2 go = ld2(D->G[cur], n);
        st2(D->G[cur ^ 1], n, make_float2(gx, gy));
        gg = (double)gx * (double)gx + (double)gy * (double)gy;
        gnum = (double)gx * ((double)gx - (double)go.x) + (double)gy * ((double)gy - (double)go.y);
    }
    gg = wave_sum(gg);
    gnum = wave_sum(gnum);
    chunk_done(gg, gnum, D->NCP, c, D->nchunks, acc, p, lane);
}

__global__ __launch_bounds__(kBlock) void k_cg_persistent(const ProbDesc* __restrict__ descs, int P,
                                                          ProbState* __restrict__ S, Accum* __restrict__ acc,
                                                          Barrier* __restrict__ bar, int max_steps,
                                                          uint64_t timeout_ticks) {
    __shared__ int s_flags[kMaxProblems];
    __shared__ float s_alpha[kMaxProblems];
    __shared__ float s_beta[kMaxProblems];
    __shared__ int s_itemA[kMaxProblems + 1];
    __shared__ int s_itemB[kMaxProblems + 1];
    __shared__ int s_any;

    if (threadIdx.x == 0) {
        int a = 0, b = 0;
        for (int p = 0; p < P; ++p) {
            s_itemA[p] = a;
            s_itemB[p] = b;
            a += descs[p].nchunks + descs[p].ntiles;
            b += descs[p].nchunks;
        }
        s_itemA[P] = a;
        s_itemB[P] = b;
    }
    __syncthreads();
    const int nA = s_itemA[P], nB = s_itemB[P];
    const int lane = threadIdx.x & (kWave - 1);
    const int gwave = __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + threadIdx.x / kWave);
    const int nwaves = gridDim.x * kWavesPerBlock;
    const bool writer = blockIdx.x == 0;

    for (int step = 0; step < max_steps; ++step) {
        const bool isA = (step & 1) == 0;
        const ProbState* Sin = S + (size_t)(step & 1) * P;
        ProbState* Sout = S + (size_t)((step + 1) & 1) * P;
        const Accum* aprev = acc + (step + 2) % 3;
        Accum* acur = acc + step % 3;
        Accum* azero = acc + (step + 1) % 3;
        if (threadIdx.x == 0) s_any = 0;
        __syncthreads();

        // ---- decisions (identical in every block; block 0 records them) ----
        for (int p = threadIdx.x; p < P; p += kBlock) {
            const ProbDesc& D = descs[p];
            ProbState st = Sin[p];
            int flags = 0;
            if (isA) {
                if (st.mode == MODE_NODE) {
                    const double gg = aprev->fin[p][0], gnum = aprev->fin[p][1];
                    st.cur ^= 1;
                    st.dcur ^= 1;
                    const float gnorm = (float)sqrt(gg);
                    if (st.k >= D.max_iter) {
                        // while_loop exits on k < max_iter before evaluating anything
                        st.mode = MODE_DONE;
                        if (writer) *D.n_iters = st.k;
                    } else if (gnorm < D.tol) {
                        st.mode = MODE_DONE;
                        if (writer) {
                            if (D.hist) {
                                D.hist[3 * st.k] = (float)st.J;
                                D.hist[3 * st.k + 1] = gnorm;
                                D.hist[3 * st.k + 2] = 0.f;
                            }
                            *D.n_iters = st.k;
                        }
                    } else {
                        float b = 0.f;
                        if (st.k > 0) {
                            b = (float)gnum / ((float)st.gg + 1e-16f);
                            if (b < 0.f) b = 0.f;  // max(beta, 0); NaN propagates like jnp.maximum
                        }
                        st.beta = b;
                        st.gg = gg;
                        st.gnorm = gnorm;
                        st.mode = MODE_PASS;
                        st.alpha = 1.f;
                        st.trial = 0;
                        st.restart = 0;
                        st.final_eval = 0;
                    }
                }
                if (st.mode == MODE_PASS || st.mode == MODE_INIT) {
                    flags |= F_DOA;
                    st.passes += 1;
                }
            } else {
                if (st.mode == MODE_PASS || st.mode == MODE_INIT) {
                    const double sr2 = (double)aprev->hi[p] + (double)aprev->lo[p] / kFixScale;
                    const double gd = aprev->fin[p][0], E = aprev->fin[p][1];
                    double Jc = 0.5 * sr2 * D.invNp + (double)D.alpha_reg * (0.25 * E);
                    if (aprev->bad[p]) Jc = NAN;
                    bool accept = false;
                    if (st.mode == MODE_INIT || st.final_eval) {
                        accept = true;
                    } else if (st.trial == 0 && !st.restart && gd >= 0.0) {
                        st.restart = 1;  // non-descent direction: redo the pass with d = -g
                    } else if (Jc <= st.J + 1e-4 * (double)st.alpha * gd) {
                        accept = true;
                    } else {
                        st.Jtrial = (float)Jc;
                        st.trial += 1;
                        st.alpha *= 0.5f;
                        if (st.trial == 10) st.final_eval = 1;  // alpha = 2^-10 is taken untested
                    }
                    if (accept) {
                        flags |= F_DOB;
                        if (st.mode != MODE_INIT) {
                            if (writer && D.hist) {
                                D.hist[3 * st.k] = st.final_eval ? st.Jtrial : (float)Jc;
                                D.hist[3 * st.k + 1] = st.gnorm;
                                D.hist[3 * st.k + 2] = st.alpha;
                            }
                            st.k += 1;
                        }
                        st.J = Jc;
                        st.mode = MODE_NODE;
                    }
                }
            }
            if (st.restart) flags |= F_RESTART;
            if (st.cur) flags |= F_CUR;
            if (st.dcur) flags |= F_DCUR;
            s_flags[p] = flags;
            s_alpha[p] = st.alpha;
            s_beta[p] = st.beta;
            if (flags & (F_DOA | F_DOB)) atomicOr(&s_any, 1);
            if (writer) {
                Sout[p] = st;
                if (st.mode == MODE_DONE) S[2 * (size_t)P + p] = st;
                azero->fin[p][0] = 0.0;
                azero->fin[p][1] = 0.0;
                azero->hi[p] = 0ull;
                azero->lo[p] = 0ull;
                azero->cnt[p] = 0u;
                azero->bad[p] = 0;
            }
        }
        __syncthreads();
        if (!s_any) break;

        // ---- work ----
        if (isA) {
            unsigned long long hi = 0ull, lo = 0ull;
            int bad = 0, accp = -1;
            for (int it = gwave; it < nA; it += nwaves) {
                const int p = find_prob(s_itemA, P, it);
                const int flags = s_flags[p];
                if (!(flags & F_DOA)) continue;
                const ProbDesc* D = descs + p;
                const int l = it - s_itemA[p];
                const int nch = D->nchunks;
                if (l < nch) {
                    node_pass_a(D, l, lane, flags, s_alpha[p], s_beta[p], acur, p);
                } else {
                    if (p != accp) {
                        if (accp >= 0 && lane == 0) {
                            if (hi) atomicAdd(&acur->hi[accp], hi);
                            if (lo) atomicAdd(&acur->lo[accp], lo);
                            if (bad) atomicOr(&acur->bad[accp], 1);
                        }
                        hi = lo = 0ull;
                        bad = 0;
                        accp = p;
                    }
                    tile_pass(D, l - nch, lane, flags, s_alpha[p], s_beta[p], hi, lo, bad);
                }
            }
            if (accp >= 0 && lane == 0) {
                if (hi) atomicAdd(&acur->hi[accp], hi);
                if (lo) atomicAdd(&acur->lo[accp], lo);
                if (bad) atomicOr(&acur->bad[accp], 1);
            }
        } else {
            for (int it = gwave; it < nB; it += nwaves) {
                const int p = find_prob(s_itemB, P, it);
                const int flags = s_flags[p];
                if (!(flags & F_DOB)) continue;
                node_pass_b(descs + p, it - s_itemB[p], lane, flags, s_alpha[p], acur, p);
            }
        }
        if (!grid_sync(bar, gridDim.x, timeout_ticks)) break;
    }
}

// Per-problem start state, zeroed g/d buffers, NaN-filled history.
__global__ void k_init(const ProbDesc* __restrict__ descs, int P, ProbState* __restrict__ S) {
    const int p = blockIdx.y;
    if (p >= P) return;
    const ProbDesc& D = descs[p];
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < 2 * D.nnodes; n += gridDim.x * blockDim.x) {
        D.G[0][n] = 0.f;
        D.D[0][n] = 0.f;
    }
    if (D.hist)
        for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < 3 * D.max_iter; k += gridDim.x * blockDim.x)
            D.hist[k] = NAN;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ProbState st;
        st.mode = MODE_INIT;
        st.k = st.trial = st.restart = st.final_eval = st.cur = st.dcur = st.passes = 0;
        st.alpha = 0.f;
        st.beta = 0.f;
        st.gnorm = 0.f;
        st.Jtrial = 0.f;
        st.J = 0.0;
        st.gg = 0.0;
        S[p] = st;
        S[P + p] = st;
        S[2 * P + p] = st;
        *D.n_iters = -1;
    }
}

// Copy the final iterate into the caller's buffer (U[0]) and publish counters;
// n_iters = -1 marks a problem that did not finish (barrier timeout).
__global__ void k_finalize(const ProbDesc* __restrict__ descs, int P, const ProbState* __restrict__ S,
                           const Barrier* __restrict__ bar) {
    const int p = blockIdx.y;
    if (p >= P) return;
    const ProbDesc& D = descs[p];
    const ProbState st = S[2 * (size_t)P + p];  // state recorded when the problem finished
    if (st.mode == MODE_DONE && st.cur == 1)
        for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < 2 * D.nnodes; n += gridDim.x * blockDim.x)
            D.U[0][n] = D.U[1][n];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (D.pass_count) *D.pass_count = st.passes;
        if (st.mode != MODE_DONE || bar->err) *D.n_iters = -1;
    }
}

// f(x_p) sampled once per reference image.
__global__ void k_sample_ref(const float* __restrict__ img, int H, int W, int interp, const float* __restrict__ px,
                             const float* __restrict__ py, int64_t n, float* __restrict__ F) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float v, gx, gy;
    sample(interp, img, H, W, px[s], py[s], v, gx, gy);
    F[s] = v;
}

// ---------------------------------------------------------------------------
// plan construction kernels
// ---------------------------------------------------------------------------
__global__ void k_validate(const int32_t* __restrict__ pnodes, int64_t n4, int64_t nn, const int32_t* __restrict__ nbi,
                           const int32_t* __restrict__ nbd, int md, int* __restrict__ err) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n4) {
        int v = pnodes[t];
        if (v < 0 || v >= nn) atomicOr(err, 1);
    }
    if (t < nn) {
        int deg = nbd[t];
        if (deg < 0) atomicOr(err, 2);
        for (int k = 0; k < md && k < deg; ++k) {
            int j = nbi[t * md + k];
            if (j < 0 || j >= nn) atomicOr(err, 2);
        }
    }
}

__global__ void k_group_keys(const int32_t* __restrict__ pelts, const int32_t* __restrict__ pnodes, int64_t np,
                             unsigned long long* __restrict__ keys, int32_t* __restrict__ vals) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    unsigned long long k;
    if (pelts)
        k = (unsigned long long)(uint32_t)pelts[p];
    else
        k = ((unsigned long long)(uint32_t)pnodes[4 * p] << 42) | ((unsigned long long)(uint32_t)pnodes[4 * p + 1] << 21) |
            (unsigned long long)(uint32_t)pnodes[4 * p + 2];
    keys[p] = k;
    vals[p] = (int32_t)p;
}

__global__ void k_heads(const unsigned long long* __restrict__ k, int64_t n, uint8_t* __restrict__ head) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) head[s] = (s == 0 || k[s] != k[s - 1]) ? 1 : 0;
}

__global__ void k_group_ntiles(const int32_t* __restrict__ gstart, int64_t ng, int64_t np, int32_t* __restrict__ nt) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    int64_t e = (g + 1 < ng) ? gstart[g + 1] : np;
    nt[g] = (int32_t)((e - gstart[g] + kTile - 1) / kTile);
}

__global__ void k_tiles(const int32_t* __restrict__ gstart, const int32_t* __restrict__ tstart, int64_t ng, int64_t np,
                        const int32_t* __restrict__ perm, const int32_t* __restrict__ pnodes, int4* __restrict__ hdr,
                        int4* __restrict__ tn) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    int s0 = gstart[g];
    int s1 = (g + 1 < ng) ? gstart[g + 1] : (int)np;
    int p0 = perm[s0];
    int4 nodes = make_int4(pnodes[4 * (int64_t)p0], pnodes[4 * (int64_t)p0 + 1], pnodes[4 * (int64_t)p0 + 2],
                           pnodes[4 * (int64_t)p0 + 3]);
    int t = tstart[g];
    for (int s = s0; s < s1; s += kTile, ++t) {
        hdr[t] = make_int4(s, min(kTile, s1 - s), (int)g, 0);
        tn[t] = nodes;
    }
}

__global__ void k_soa(const int32_t* __restrict__ perm, int64_t np, const float* __restrict__ coords,
                      const float* __restrict__ shapeN, float* __restrict__ px, float* __restrict__ py,
                      float4* __restrict__ N4) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= np) return;
    int64_t p = perm[s];
    px[s] = coords[2 * p];
    py[s] = coords[2 * p + 1];
    N4[s] = make_float4(shapeN[4 * p], shapeN[4 * p + 1], shapeN[4 * p + 2], shapeN[4 * p + 3]);
}

__global__ void k_tile_entries(const int4* __restrict__ tn, int64_t nt, int32_t* __restrict__ keys,
                               int32_t* __restrict__ vals, int32_t* __restrict__ cnt) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 4 * nt) return;
    int4 n = tn[e / 4];
    int a = (int)(e % 4);
    int node = a == 0 ? n.x : (a == 1 ? n.y : (a == 2 ? n.z : n.w));
    keys[e] = node;
    vals[e] = (int32_t)e;
    atomicAdd(&cnt[node], 1);
}

__global__ void k_reg_table(const int32_t* __restrict__ nbi, const int32_t* __restrict__ nbd,
                            const float* __restrict__ nbw, const float* __restrict__ regw, int64_t nn, int md,
                            int32_t* __restrict__ oi, float* __restrict__ ow) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn * md) return;
    int64_t n = t / md;
    int k = (int)(t % md);
    int j = nbi[t];
    if (k < nbd[n] && j >= 0) {
        oi[t] = j;
        // reg_energy_spring_global: w * 0.5 * (w_i + w_j) (solver_global_cg.py:279-280)
        ow[t] = nbw[t] * 0.5f * (regw[n] + regw[j]);
    } else {
        oi[t] = -1;
        ow[t] = 0.f;
    }
}

static int bits_for(int64_t n) {
    int b = 1;
    while ((int64_t(1) << b) < n) ++b;
    return b;
}

}  // namespace d2ic

using namespace d2ic;

struct d2ic_cg_plan {
    int device = 0;
    int64_t np = 0, nn = 0, ntiles = 0, nchunks = 0;
    int nb_maxdeg = 0;
    int4* tile_hdr = nullptr;
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
    double* NCP = nullptr;
    int ref_H = 0, ref_W = 0, interp = 0;
    bool has_ref = false;
    size_t bytes = 0;
    std::vector<void*> allocs;
};

namespace {

template <class T>
hipError_t plan_alloc(d2ic_cg_plan* pl, T** out, size_t count) {
    size_t bytes = sizeof(T) * (count ? count : 1);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return e;
    pl->allocs.push_back(p);
    pl->bytes += bytes;
    *out = reinterpret_cast<T*>(p);
    return hipSuccess;
}

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

// Pinned staging for per-launch descriptors (thread-local, grow-only).
struct Staging {
    void* host = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
};
thread_local Staging g_stage;

int occupancy_blocks(int* out) {
    static std::mutex mu;
    static int cached_dev = -1, cached = 0;
    int dev = 0;
    D2IC_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    if (dev != cached_dev) {
        int per_cu = 0;
        D2IC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cg_persistent, kBlock, 0));
        hipDeviceProp_t prop;
        D2IC_HIP(hipGetDeviceProperties(&prop, dev));
        int coop = 0;
        D2IC_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
        if (!coop) {
            set_error("device %d does not support cooperative launches", dev);
            return D2IC_ERR_DEVICE;
        }
        cached = std::max(1, per_cu) * prop.multiProcessorCount;
        cached_dev = dev;
    }
    *out = cached;
    return D2IC_OK;
}

}  // namespace

extern "C" {

int d2ic_cg_plan_create(const float* pixel_coords, const int32_t* pixel_nodes, const float* pixel_shapeN,
                        const int32_t* pixel_elements, int64_t n_pixels, int64_t n_nodes, const int32_t* nb_index,
                        const int32_t* nb_degree, const float* nb_weight, const float* reg_weight,
                        int32_t nb_max_degree, d2ic_cg_plan** plan_out, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    D2IC_CHECK_ARG(plan_out != nullptr, "plan must not be NULL");
    D2IC_CHECK_ARG(n_pixels >= 0 && n_nodes >= 0 && nb_max_degree >= 0, "sizes must be non-negative");
    D2IC_CHECK_ARG(n_nodes < (1 << 21), "at most 2^21 - 1 nodes are supported");
    D2IC_CHECK_ARG(4 * n_pixels < (int64_t)INT32_MAX, "too many pixels");
    *plan_out = nullptr;
    d2ic_cg_plan* pl = new d2ic_cg_plan();
    auto fail = [&](int code) {
        d2ic_cg_plan_destroy(pl);
        return code;
    };
#define PA(ptr, n)                                                                 \
    do {                                                                           \
        if (plan_alloc(pl, &(ptr), (n)) != hipSuccess) {                           \
            set_error("device allocation of %s failed", #ptr);                     \
            return fail(D2IC_ERR_NOMEM);                                           \
        }                                                                          \
    } while (0)
#define PH(call)                                                                   \
    do {                                                                           \
        hipError_t _e = (call);                                                    \
        if (_e != hipSuccess) {                                                    \
            set_error("%s failed: %s", #call, hipGetErrorString(_e));              \
            return fail(D2IC_ERR_DEVICE);                                          \
        }                                                                          \
    } while (0)
    PH(hipGetDevice(&pl->device));
    pl->np = n_pixels;
    pl->nn = n_nodes;
    pl->nb_maxdeg = nb_max_degree;
    pl->nchunks = ceil_div(n_nodes, kWave);
    const int tb = 256;

    // validate indices once (a bad index would otherwise be read out of bounds)
    {
        Tmp err;
        PH(err.alloc(sizeof(int), st));
        PH(hipMemsetAsync(err.p, 0, sizeof(int), st));
        int64_t nmax = std::max<int64_t>(4 * n_pixels, n_nodes);
        if (nmax > 0) {
            k_validate<<<ceil_div(nmax, tb), tb, 0, st>>>(pixel_nodes, 4 * n_pixels, n_nodes, nb_index, nb_degree,
                                                          nb_max_degree, err.as<int>());
            PH(hipGetLastError());
        }
        int herr = 0;
        PH(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
        PH(hipStreamSynchronize(st));
        if (herr & 1) {
            set_error("pixel_nodes holds an index outside [0, %lld)", (long long)n_nodes);
            return fail(D2IC_ERR_INVALID);
        }
        if (herr & 2) {
            set_error("node_neighbor_index/degree hold an invalid entry");
            return fail(D2IC_ERR_INVALID);
        }
    }

    // node arrays
    PA(pl->U1, 2 * n_nodes);
    PA(pl->G[0], 2 * n_nodes);
    PA(pl->G[1], 2 * n_nodes);
    PA(pl->D[0], 2 * n_nodes);
    PA(pl->D[1], 2 * n_nodes);
    PA(pl->NCP, 2 * pl->nchunks);
    PA(pl->node_ptr, n_nodes + 1);
    PA(pl->nb_idx, n_nodes * nb_max_degree);
    PA(pl->nb_wp, n_nodes * nb_max_degree);
    PA(pl->px, n_pixels);
    PA(pl->py, n_pixels);
    PA(pl->N4, n_pixels);
    PA(pl->F, n_pixels);
    if (n_nodes * nb_max_degree > 0) {
        k_reg_table<<<ceil_div(n_nodes * nb_max_degree, tb), tb, 0, st>>>(nb_index, nb_degree, nb_weight, reg_weight,
                                                                          n_nodes, nb_max_degree, pl->nb_idx,
                                                                          pl->nb_wp);
        PH(hipGetLastError());
    }
    PH(hipMemsetAsync(pl->node_ptr, 0, sizeof(int32_t) * (n_nodes + 1), st));

    if (n_pixels > 0) {
        // group pixels by element (stable)
        Tmp keys, keys_s, vals, perm, head, gstart, nsel, ntg, tstart, tmp;
        PH(keys.alloc(sizeof(unsigned long long) * n_pixels, st));
        PH(keys_s.alloc(sizeof(unsigned long long) * n_pixels, st));
        PH(vals.alloc(sizeof(int32_t) * n_pixels, st));
        PH(perm.alloc(sizeof(int32_t) * n_pixels, st));
        PH(head.alloc(n_pixels, st));
        PH(gstart.alloc(sizeof(int32_t) * n_pixels, st));
        PH(nsel.alloc(sizeof(int32_t), st));
        k_group_keys<<<ceil_div(n_pixels, tb), tb, 0, st>>>(pixel_elements, pixel_nodes, n_pixels,
                                                            keys.as<unsigned long long>(), vals.as<int32_t>());
        PH(hipGetLastError());
        size_t b1 = 0, b2 = 0, b3 = 0;
        PH(hipcub::DeviceRadixSort::SortPairs(nullptr, b1, keys.as<unsigned long long>(),
                                              keys_s.as<unsigned long long>(), vals.as<int32_t>(),
                                              perm.as<int32_t>(), (int)n_pixels, 0, 64, st));
        hipcub::CountingInputIterator<int32_t> cnt_it(0);
        PH(hipcub::DeviceSelect::Flagged(nullptr, b2, cnt_it, head.as<uint8_t>(), gstart.as<int32_t>(),
                                         nsel.as<int32_t>(), (int)n_pixels, st));
        PH(hipcub::DeviceScan::ExclusiveSum(nullptr, b3, gstart.as<int32_t>(), gstart.as<int32_t>(), (int)n_pixels,
                                            st));
        PH(tmp.alloc(std::max(b1, std::max(b2, b3)), st));
        PH(hipcub::DeviceRadixSort::SortPairs(tmp.p, b1, keys.as<unsigned long long>(),
                                              keys_s.as<unsigned long long>(), vals.as<int32_t>(),
                                              perm.as<int32_t>(), (int)n_pixels, 0, 64, st));
        k_heads<<<ceil_div(n_pixels, tb), tb, 0, st>>>(keys_s.as<unsigned long long>(), n_pixels,
                                                       head.as<uint8_t>());
        PH(hipGetLastError());
        PH(hipcub::DeviceSelect::Flagged(tmp.p, b2, cnt_it, head.as<uint8_t>(), gstart.as<int32_t>(),
                                         nsel.as<int32_t>(), (int)n_pixels, st));
        int32_t ng = 0;
        PH(hipMemcpyAsync(&ng, nsel.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        PH(hipStreamSynchronize(st));
        PH(ntg.alloc(sizeof(int32_t) * (ng + 1), st));
        PH(tstart.alloc(sizeof(int32_t) * (ng + 1), st));
        k_group_ntiles<<<ceil_div(ng, tb), tb, 0, st>>>(gstart.as<int32_t>(), ng, n_pixels, ntg.as<int32_t>());
        PH(hipGetLastError());
        PH(hipMemsetAsync(ntg.as<int32_t>() + ng, 0, sizeof(int32_t), st));
        size_t b4 = 0;
        PH(hipcub::DeviceScan::ExclusiveSum(nullptr, b4, ntg.as<int32_t>(), tstart.as<int32_t>(), ng + 1, st));
        Tmp tmp2;
        PH(tmp2.alloc(b4, st));
        PH(hipcub::DeviceScan::ExclusiveSum(tmp2.p, b4, ntg.as<int32_t>(), tstart.as<int32_t>(), ng + 1, st));
        int32_t nt = 0;
        PH(hipMemcpyAsync(&nt, tstart.as<int32_t>() + ng, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        PH(hipStreamSynchronize(st));
        pl->ntiles = nt;
        PA(pl->tile_hdr, nt);
        PA(pl->tile_nodes, nt);
        PA(pl->TP, 8 * (size_t)nt);
        PA(pl->node_ent, 4 * (size_t)nt);
        k_tiles<<<ceil_div(ng, tb), tb, 0, st>>>(gstart.as<int32_t>(), tstart.as<int32_t>(), ng, n_pixels,
                                                 perm.as<int32_t>(), pixel_nodes, pl->tile_hdr, pl->tile_nodes);
        PH(hipGetLastError());
        k_soa<<<ceil_div(n_pixels, tb), tb, 0, st>>>(perm.as<int32_t>(), n_pixels, pixel_coords, pixel_shapeN, pl->px,
                                                     pl->py, pl->N4);
        PH(hipGetLastError());

        // node -> (tile, local node) gather table
        int64_t ne = 4 * (int64_t)nt;
        Tmp ek, ev, eks, cnt, tmp3;
        PH(ek.alloc(sizeof(int32_t) * ne, st));
        PH(ev.alloc(sizeof(int32_t) * ne, st));
        PH(eks.alloc(sizeof(int32_t) * ne, st));
        PH(cnt.alloc(sizeof(int32_t) * (n_nodes + 1), st));
        PH(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (n_nodes + 1), st));
        k_tile_entries<<<ceil_div(ne, tb), tb, 0, st>>>(pl->tile_nodes, nt, ek.as<int32_t>(), ev.as<int32_t>(),
                                                        cnt.as<int32_t>());
        PH(hipGetLastError());
        size_t b5 = 0, b6 = 0;
        int bits = bits_for(n_nodes);
        PH(hipcub::DeviceRadixSort::SortPairs(nullptr, b5, ek.as<int32_t>(), eks.as<int32_t>(), ev.as<int32_t>(),
                                              pl->node_ent, (int)ne, 0, bits, st));
        PH(hipcub::DeviceScan::ExclusiveSum(nullptr, b6, cnt.as<int32_t>(), pl->node_ptr, (int)(n_nodes + 1), st));
        PH(tmp3.alloc(std::max(b5, b6), st));
        PH(hipcub::DeviceRadixSort::SortPairs(tmp3.p, b5, ek.as<int32_t>(), eks.as<int32_t>(), ev.as<int32_t>(),
                                              pl->node_ent, (int)ne, 0, bits, st));
        PH(hipcub::DeviceScan::ExclusiveSum(tmp3.p, b6, cnt.as<int32_t>(), pl->node_ptr, (int)(n_nodes + 1), st));
        PH(hipStreamSynchronize(st));
    }
#undef PA
#undef PH
    *plan_out = pl;
    return D2IC_OK;
}

int d2ic_cg_plan_destroy(d2ic_cg_plan* plan) {
    if (!plan) return D2IC_OK;
    for (void* p : plan->allocs) (void)hipFree(p);
    delete plan;
    return D2IC_OK;
}

int d2ic_cg_plan_info(const d2ic_cg_plan* plan, int64_t* n_pixels, int64_t* n_nodes, int64_t* n_tiles,
                      int64_t* device_bytes) {
    D2IC_CHECK_ARG(plan != nullptr, "plan must not be NULL");
    if (n_pixels) *n_pixels = plan->np;
    if (n_nodes) *n_nodes = plan->nn;
    if (n_tiles) *n_tiles = plan->ntiles;
    if (device_bytes) *device_bytes = (int64_t)plan->bytes;
    return D2IC_OK;
}

int d2ic_cg_set_reference(d2ic_cg_plan* plan, const float* ref_image, int32_t H, int32_t W, int32_t interp,
                          void* stream) {
    hipStream_t st = (hipStream_t)stream;
    D2IC_CHECK_ARG(plan != nullptr, "plan must not be NULL");
    D2IC_CHECK_ARG(ref_image != nullptr && H > 0 && W > 0, "reference image must be non-empty");
    D2IC_CHECK_ARG(interp == D2IC_INTERP_CUBIC || interp == D2IC_INTERP_LINEAR, "unsupported interpolation %d",
                   interp);
    if (plan->np > 0) {
        k_sample_ref<<<ceil_div(plan->np, 256), 256, 0, st>>>(ref_image, H, W, interp, plan->px, plan->py, plan->np,
                                                              plan->F);
        D2IC_LAUNCH_CHECK();
    }
    plan->ref_H = H;
    plan->ref_W = W;
    plan->interp = interp;
    plan->has_ref = true;
    return D2IC_OK;
}

int d2ic_cg_solve_batch(int32_t n_problems, d2ic_cg_plan* const* plans, const float* const* def_images,
                        const int32_t* H, const int32_t* W, float* const* u, const float* alpha_reg,
                        const int32_t* max_iter, const float* tol, float* const* hist, int32_t* const* n_iters,
                        int32_t* const* pass_count, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    const int P = n_problems;
    D2IC_CHECK_ARG(P >= 1 && P <= kMaxProblems, "n_problems must be in [1, %d]", kMaxProblems);
    D2IC_CHECK_ARG(plans && def_images && H && W && u && alpha_reg && max_iter && tol && n_iters,
                   "batch arrays must not be NULL");
    std::vector<ProbDesc> hd(P);
    int max_it = 0;
    int64_t total_items = 0;
    for (int p = 0; p < P; ++p) {
        d2ic_cg_plan* pl = plans[p];
        D2IC_CHECK_ARG(pl != nullptr, "plan %d is NULL", p);
        if (!pl->has_ref) {
            set_error("plan %d has no reference image (call d2ic_cg_set_reference first)", p);
            return D2IC_ERR_STATE;
        }
        for (int q = 0; q < p; ++q) D2IC_CHECK_ARG(plans[q] != pl, "plan %d appears twice in the batch", p);
        D2IC_CHECK_ARG(def_images[p] != nullptr && H[p] > 0 && W[p] > 0, "deformed image %d must be non-empty", p);
        D2IC_CHECK_ARG(u[p] != nullptr || pl->nn == 0, "u[%d] must not be NULL", p);
        D2IC_CHECK_ARG(n_iters[p] != nullptr, "n_iters[%d] must not be NULL", p);
        D2IC_CHECK_ARG(max_iter[p] >= 0 && max_iter[p] <= (1 << 20), "max_iter[%d] out of range", p);
        ProbDesc& d = hd[p];
        d.tile_hdr = pl->tile_hdr;
        d.tile_nodes = pl->tile_nodes;
        d.px = pl->px;
        d.py = pl->py;
        d.N4 = pl->N4;
        d.F = pl->F;
        d.node_ptr = pl->node_ptr;
        d.node_ent = pl->node_ent;
        d.nb_idx = pl->nb_idx;
        d.nb_wp = pl->nb_wp;
        d.U[0] = u[p];
        d.U[1] = pl->U1;
        d.G[0] = pl->G[0];
        d.G[1] = pl->G[1];
        d.D[0] = pl->D[0];
        d.D[1] = pl->D[1];
        d.TP = pl->TP;
        d.NCP = pl->NCP;
        d.img = def_images[p];
        d.hist = hist ? hist[p] : nullptr;
        d.n_iters = n_iters[p];
        d.pass_count = pass_count ? pass_count[p] : nullptr;
        d.H = H[p];
        d.W = W[p];
        d.interp = pl->interp;
        d.nb_maxdeg = pl->nb_maxdeg;
        d.ntiles = (int)pl->ntiles;
        d.nnodes = (int)pl->nn;
        d.nchunks = (int)pl->nchunks;
        d.max_iter = max_iter[p];
        d.alpha_reg = alpha_reg[p];
        d.tol = tol[p];
        d.invNp = 1.0 / (double)(pl->np > 0 ? pl->np : 1);
        max_it = std::max(max_it, max_iter[p]);
        total_items += pl->ntiles + pl->nchunks;
    }
    D2IC_CHECK_ARG(total_items < (int64_t)INT32_MAX, "batch too large");

    const size_t off_S = ((sizeof(ProbDesc) * P + 255) / 256) * 256;
    const size_t off_A = off_S + ((sizeof(ProbState) * 3 * P + 255) / 256) * 256;
    const size_t off_B = off_A + ((sizeof(Accum) * 3 + 255) / 256) * 256;
    const size_t bytes = off_B + 256;
    Tmp ws;
    D2IC_HIP(ws.alloc(bytes, st));
    char* base = ws.as<char>();
    ProbDesc* d_desc = reinterpret_cast<ProbDesc*>(base);
    ProbState* d_S = reinterpret_cast<ProbState*>(base + off_S);
    Accum* d_acc = reinterpret_cast<Accum*>(base + off_A);
    Barrier* d_bar = reinterpret_cast<Barrier*>(base + off_B);

    Staging& sg = g_stage;
    if (sg.pending) {
        D2IC_HIP(hipEventSynchronize(sg.ev));
        sg.pending = false;
    }
    if (sg.cap < sizeof(ProbDesc) * P) {
        if (sg.host) (void)hipHostFree(sg.host);
        sg.host = nullptr;
        sg.cap = 0;
        D2IC_HIP(hipHostMalloc(&sg.host, sizeof(ProbDesc) * kMaxProblems, hipHostMallocDefault));
        sg.cap = sizeof(ProbDesc) * kMaxProblems;
    }
    if (!sg.ev) D2IC_HIP(hipEventCreateWithFlags(&sg.ev, hipEventDisableTiming));
    memcpy(sg.host, hd.data(), sizeof(ProbDesc) * P);
    D2IC_HIP(hipMemcpyAsync(d_desc, sg.host, sizeof(ProbDesc) * P, hipMemcpyHostToDevice, st));
    D2IC_HIP(hipEventRecord(sg.ev, st));
    sg.pending = true;
    D2IC_HIP(hipMemsetAsync(d_acc, 0, sizeof(Accum) * 3, st));
    D2IC_HIP(hipMemsetAsync(d_bar, 0, sizeof(Barrier), st));

    int maxnn = 1;
    for (int p = 0; p < P; ++p) maxnn = std::max<int>(maxnn, (int)std::max<int64_t>(plans[p]->nn * 2, 3 * max_iter[p]));
    dim3 gi((unsigned)std::min<int64_t>(ceil_div(maxnn, 256), 1024), (unsigned)P);
    k_init<<<gi, 256, 0, st>>>(d_desc, P, d_S);
    D2IC_LAUNCH_CHECK();

    int coop_max = 0;
    int rc = occupancy_blocks(&coop_max);
    if (rc != D2IC_OK) return rc;
    int64_t want = ceil_div(total_items, (int64_t)kWavesPerBlock * 2);
    int nblocks = (int)std::max<int64_t>(1, std::min<int64_t>(coop_max, want));
    int max_steps = 8 + 2 * (2 + max_it * 12);
    uint64_t timeout_ticks = 100000000ull * 4;  // wall_clock64 runs at 100 MHz: 4 s per barrier wait
    void* args[] = {(void*)&d_desc, (void*)&P, (void*)&d_S, (void*)&d_acc, (void*)&d_bar, (void*)&max_steps,
                    (void*)&timeout_ticks};
    D2IC_HIP(hipLaunchCooperativeKernel((const void*)k_cg_persistent, dim3(nblocks), dim3(kBlock), args, 0, st));
    k_finalize<<<gi, 256, 0, st>>>(d_desc, P, d_S, d_bar);
    D2IC_LAUNCH_CHECK();
    return D2IC_OK;
}

int d2ic_cg_solve(d2ic_cg_plan* plan, const float* def_image, int32_t H, int32_t W, float* u, float alpha_reg,
                  int32_t max_iter, float tol, float* hist, int32_t* n_iters, int32_t* pass_count, void* stream) {
    d2ic_cg_plan* plans[1] = {plan};
    const float* imgs[1] = {def_image};
    float* us[1] = {u};
    float* hs[1] = {hist};
    int32_t* ns[1] = {n_iters};
    int32_t* pc[1] = {pass_count};
    return d2ic_cg_solve_batch(1, plans, imgs, &H, &W, us, &alpha_reg, &max_iter, &tol, hs, ns, pc, stream);
}

}  // extern "C"
