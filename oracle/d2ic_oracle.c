/*
 * d2ic_oracle.c -- CPU restatement of the D2IC global FE-DIC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the library
 * built from it.  The product (d2ic_amd) never links or calls it.
 *
 * The file is compiled twice, once with REAL=float (stands in for the
 * reference's default JAX FP32 run) and once with REAL=double (stands in for
 * the reference's x64 run).  Every exported symbol carries the suffix SFX
 * (_f32 / _f64).  Built with -ffp-contract=off so that every product and sum
 * is rounded as written.
 *
 * Each function names the reference source it restates (paths relative to
 * the reference repository root).  Reductions over pixels/nodes accumulate in
 * double in both instances and are combined in a fixed order, so results do
 * not depend on the OpenMP thread count.
 *
 * Pinning status (what checks each function) is listed in the header of
 * oracle/oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL float
#define SFX f32
#endif

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

typedef int64_t i64;
typedef int32_t i32;

/* ------------------------------------------------------------------------ */
/* Q4 shape functions (pixel_assets.py:297-327)                               */
/* ------------------------------------------------------------------------ */
static void shape_q4(REAL xi, REAL eta, REAL N[4]) {
    const REAL q = (REAL)0.25;
    N[0] = q * ((REAL)1 - xi) * ((REAL)1 - eta);
    N[1] = q * ((REAL)1 + xi) * ((REAL)1 - eta);
    N[2] = q * ((REAL)1 + xi) * ((REAL)1 + eta);
    N[3] = q * ((REAL)1 - xi) * ((REAL)1 + eta);
}

static void shape_q4_grad(REAL xi, REAL eta, REAL dxi[4], REAL deta[4]) {
    const REAL q = (REAL)0.25;
    dxi[0] = -q * ((REAL)1 - eta);
    dxi[1] = q * ((REAL)1 - eta);
    dxi[2] = q * ((REAL)1 + eta);
    dxi[3] = -q * ((REAL)1 + eta);
    deta[0] = -q * ((REAL)1 - xi);
    deta[1] = -q * ((REAL)1 + xi);
    deta[2] = q * ((REAL)1 + xi);
    deta[3] = q * ((REAL)1 - xi);
}

/* 2x2 solve A x = f by LU with partial pivoting (the getrf + two trsm
 * sequence jnp.linalg.solve lowers to; pixel_assets.py:290, strain.py:81-82).
 * A = [[a, b], [c, d]]. */
static void solve2(REAL a, REAL b, REAL c, REAL d, REAL f0, REAL f1, REAL* x0, REAL* x1) {
    if (fabs((double)c) > fabs((double)a)) {
        REAL t;
        t = a; a = c; c = t;
        t = b; b = d; d = t;
        t = f0; f0 = f1; f1 = t;
    }
    REAL l = c / a;
    REAL u22 = d - l * b;
    REAL y1 = f1 - l * f0;
    REAL s1 = y1 / u22;
    REAL s0 = (f0 - b * s1) / a;
    *x0 = s0;
    *x1 = s1;
}

/* _newton_quadrature (pixel_assets.py:277-294): 10 Newton steps from (0,0). */
static void newton_q4(REAL px, REAL py, const REAL X[4], const REAL Y[4], REAL* xi_out, REAL* eta_out) {
    REAL xi = 0, eta = 0;
    for (int it = 0; it < 10; ++it) {
        REAL N[4], dxi[4], deta[4];
        shape_q4(xi, eta, N);
        shape_q4_grad(xi, eta, dxi, deta);
        REAL x = 0, y = 0, j00 = 0, j01 = 0, j10 = 0, j11 = 0;
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
        REAL F0 = x - px, F1 = y - py, d0, d1;
        solve2(j00, j01, j10, j11, F0, F1, &d0, &d1);
        xi = xi - d0;
        eta = eta - d1;
    }
    *xi_out = xi;
    *eta_out = eta;
}

/* ------------------------------------------------------------------------ */
/* Pixel -> element first match (pixel_assets.py:135-237)                      */
/* ------------------------------------------------------------------------ */
typedef struct {
    REAL qx[4], qy[4];
    REAL ex[4], ey[4];
    REAL bx0, bx1, by0, by1;
    REAL orient;
} quad_t;

static void quad_setup(const REAL* nodes, const i32* el, quad_t* q) {
    for (int k = 0; k < 4; ++k) {
        q->qx[k] = nodes[2 * (i64)el[k]];
        q->qy[k] = nodes[2 * (i64)el[k] + 1];
    }
    q->bx0 = q->bx1 = q->qx[0];
    q->by0 = q->by1 = q->qy[0];
    for (int k = 1; k < 4; ++k) {
        if (q->qx[k] < q->bx0) q->bx0 = q->qx[k];
        if (q->qx[k] > q->bx1) q->bx1 = q->qx[k];
        if (q->qy[k] < q->by0) q->by0 = q->qy[k];
        if (q->qy[k] > q->by1) q->by1 = q->qy[k];
    }
    REAL area2 = 0;
    for (int k = 0; k < 4; ++k) {
        int kn = (k + 1) & 3;
        q->ex[k] = q->qx[kn] - q->qx[k];
        q->ey[k] = q->qy[kn] - q->qy[k];
        area2 = area2 + (q->qx[k] * q->qy[kn] - q->qy[k] * q->qx[kn]);
    }
    q->orient = area2 >= (REAL)0 ? (REAL)1 : (REAL)-1;
}

/* _assign_chunk's test for one (pixel, element) pair: inclusive bbox and four
 * half-planes orient*cross >= -1e-6 (pixel_assets.py:193-202). */
static int quad_contains(const quad_t* q, REAL x, REAL y) {
    if (!(x >= q->bx0 && x <= q->bx1 && y >= q->by0 && y <= q->by1)) return 0;
    const REAL tol = (REAL)1e-6;
    for (int k = 0; k < 4; ++k) {
        REAL vx = x - q->qx[k], vy = y - q->qy[k];
        REAL cr = q->ex[k] * vy - q->ey[k] * vx;
        if (!(q->orient * cr >= -tol)) return 0;
    }
    return 1;
}

/* Brute-force first match over all elements: the lowest element index that
 * passes (argmax of the boolean row, pixel_assets.py:204-206). */
i32 FN(orc_first_element)(const REAL* nodes, const i32* elems, i64 Ne, REAL x, REAL y) {
    for (i64 e = 0; e < Ne; ++e) {
        quad_t q;
        quad_setup(nodes, elems + 4 * e, &q);
        if (quad_contains(&q, x, y)) return (i32)e;
    }
    return -1;
}

/* build_pixel_to_element_mapping_numpy (pixel_assets.py:135-237).
 * pixel_xy: (n, 2).  Outputs pixel_elts (n,) with -1 outside and
 * pixel_nodes (n, 4) with zeros for outside pixels.
 * use_bins != 0 accelerates with a uniform bin grid over element bounding
 * boxes; the result is the same lowest passing index (every element whose
 * inclusive bbox holds the pixel is registered in the pixel's bin). */
int FN(orc_pixel_to_element)(const REAL* pixel_xy, i64 n, const REAL* nodes, const i32* elems, i64 Ne,
                             int use_bins, i32* pixel_elts, i32* pixel_nodes) {
    if (Ne == 0) {
        for (i64 p = 0; p < n; ++p) {
            pixel_elts[p] = -1;
            for (int a = 0; a < 4; ++a) pixel_nodes[4 * p + a] = 0;
        }
        return 0;
    }
    quad_t* quads = (quad_t*)malloc(sizeof(quad_t) * (size_t)Ne);
    if (!quads) return -1;
    double gx0 = 1e300, gy0 = 1e300, gx1 = -1e300, gy1 = -1e300;
    for (i64 e = 0; e < Ne; ++e) {
        quad_setup(nodes, elems + 4 * e, &quads[e]);
        if (quads[e].bx0 < gx0) gx0 = quads[e].bx0;
        if (quads[e].by0 < gy0) gy0 = quads[e].by0;
        if (quads[e].bx1 > gx1) gx1 = quads[e].bx1;
        if (quads[e].by1 > gy1) gy1 = quads[e].by1;
    }
    i64 *bin_start = NULL, *bin_list = NULL;
    i64 nbx = 1, nby = 1;
    const double bs = 32.0;
    if (use_bins) {
        nbx = (i64)floor((gx1 - gx0) / bs) + 1;
        nby = (i64)floor((gy1 - gy0) / bs) + 1;
        bin_start = (i64*)calloc((size_t)(nbx * nby + 1), sizeof(i64));
        i64* fill = (i64*)calloc((size_t)(nbx * nby), sizeof(i64));
        for (i64 e = 0; e < Ne; ++e) {
            i64 bx0 = (i64)floor(((double)quads[e].bx0 - gx0) / bs), bx1 = (i64)floor(((double)quads[e].bx1 - gx0) / bs);
            i64 by0 = (i64)floor(((double)quads[e].by0 - gy0) / bs), by1 = (i64)floor(((double)quads[e].by1 - gy0) / bs);
            for (i64 by = by0; by <= by1; ++by)
                for (i64 bx = bx0; bx <= bx1; ++bx) bin_start[by * nbx + bx + 1]++;
        }
        for (i64 b = 0; b < nbx * nby; ++b) bin_start[b + 1] += bin_start[b];
        bin_list = (i64*)malloc(sizeof(i64) * (size_t)(bin_start[nbx * nby] + 1));
        for (i64 e = 0; e < Ne; ++e) {
            i64 bx0 = (i64)floor(((double)quads[e].bx0 - gx0) / bs), bx1 = (i64)floor(((double)quads[e].bx1 - gx0) / bs);
            i64 by0 = (i64)floor(((double)quads[e].by0 - gy0) / bs), by1 = (i64)floor(((double)quads[e].by1 - gy0) / bs);
            for (i64 by = by0; by <= by1; ++by)
                for (i64 bx = bx0; bx <= bx1; ++bx) {
                    i64 b = by * nbx + bx;
                    bin_list[bin_start[b] + fill[b]++] = e; /* ascending e */
                }
        }
        free(fill);
    }
#pragma omp parallel for schedule(static)
    for (i64 p = 0; p < n; ++p) {
        REAL x = pixel_xy[2 * p], y = pixel_xy[2 * p + 1];
        i32 found = -1;
        if (use_bins) {
            double fx = ((double)x - gx0) / bs, fy = ((double)y - gy0) / bs;
            if (fx >= 0 && fy >= 0) {
                i64 bx = (i64)floor(fx), by = (i64)floor(fy);
                if (bx < nbx && by < nby) {
                    i64 b = by * nbx + bx;
                    for (i64 k = bin_start[b]; k < bin_start[b + 1]; ++k) {
                        i64 e = bin_list[k];
                        if (quad_contains(&quads[e], x, y)) { found = (i32)e; break; }
                    }
                }
            }
        } else {
            for (i64 e = 0; e < Ne; ++e)
                if (quad_contains(&quads[e], x, y)) { found = (i32)e; break; }
        }
        pixel_elts[p] = found;
        for (int a = 0; a < 4; ++a) pixel_nodes[4 * p + a] = found >= 0 ? elems[4 * (i64)found + a] : 0;
    }
    free(quads);
    free(bin_start);
    free(bin_list);
    return 0;
}

/* compute_pixel_shape_functions_jax (pixel_assets.py:240-274). */
int FN(orc_shape_functions)(const REAL* pixel_xy, const i32* pixel_nodes, i64 n, const REAL* nodes,
                            REAL* shapeN, REAL* xieta) {
#pragma omp parallel for schedule(static)
    for (i64 p = 0; p < n; ++p) {
        REAL X[4], Y[4];
        for (int a = 0; a < 4; ++a) {
            X[a] = nodes[2 * (i64)pixel_nodes[4 * p + a]];
            Y[a] = nodes[2 * (i64)pixel_nodes[4 * p + a] + 1];
        }
        REAL xi, eta, N[4];
        newton_q4(pixel_xy[2 * p], pixel_xy[2 * p + 1], X, Y, &xi, &eta);
        shape_q4(xi, eta, N);
        for (int a = 0; a < 4; ++a) shapeN[4 * p + a] = N[a];
        if (xieta) {
            xieta[2 * p] = xi;
            xieta[2 * p + 1] = eta;
        }
    }
    return 0;
}

/* build_pixel_assets candidate stage (pixel_assets.py:58-79, 115-132):
 * bbox sampling (row-major, i outer), ROI filter, element assignment and
 * dropping of unassigned pixels.  Buffers are sized by `cap` (>= number of
 * candidates).  Returns the number of kept pixels, or -1. */
i64 FN(orc_pixel_candidates)(const REAL* nodes, i64 Nn, const i32* elems, i64 Ne, int H, int W,
                             const uint8_t* roi, double pad, int use_bins, i64 cap,
                             REAL* coords, i32* pnodes, i32* pelts, i64* roi_flat) {
    if (Nn == 0) return 0;
    double xmin = nodes[0], xmax = nodes[0], ymin = nodes[1], ymax = nodes[1];
    for (i64 k = 1; k < Nn; ++k) {
        double x = nodes[2 * k], y = nodes[2 * k + 1];
        if (x < xmin) xmin = x;
        if (x > xmax) xmax = x;
        if (y < ymin) ymin = y;
        if (y > ymax) ymax = y;
    }
    i64 j0 = (i64)floor(xmin - pad); if (j0 < 0) j0 = 0;
    i64 j1 = (i64)ceil(xmax + pad);  if (j1 > W) j1 = W;
    i64 i0 = (i64)floor(ymin - pad); if (i0 < 0) i0 = 0;
    i64 i1 = (i64)ceil(ymax + pad);  if (i1 > H) i1 = H;
    if (j1 <= j0 || i1 <= i0) return 0;
    i64 ncand = (j1 - j0) * (i1 - i0);
    REAL* cxy = (REAL*)malloc(sizeof(REAL) * 2 * (size_t)ncand);
    i64 m = 0;
    for (i64 i = i0; i < i1; ++i)
        for (i64 j = j0; j < j1; ++j) {
            double x = (double)j + 0.5, y = (double)i + 0.5;
            if (roi) {
                i64 rr = (i64)floor(y - 0.5), cc = (i64)floor(x - 0.5);
                if (rr < 0) rr = 0; if (rr > H - 1) rr = H - 1;
                if (cc < 0) cc = 0; if (cc > W - 1) cc = W - 1;
                if (!roi[rr * W + cc]) continue;
            }
            cxy[2 * m] = (REAL)x;
            cxy[2 * m + 1] = (REAL)y;
            ++m;
        }
    i32* el = (i32*)malloc(sizeof(i32) * (size_t)(m + 1));
    i32* nd = (i32*)malloc(sizeof(i32) * 4 * (size_t)(m + 1));
    FN(orc_pixel_to_element)(cxy, m, nodes, elems, Ne, use_bins, el, nd);
    i64 k = 0;
    for (i64 p = 0; p < m; ++p) {
        if (el[p] < 0) continue;
        if (k >= cap) { k = -1; break; }
        coords[2 * k] = cxy[2 * p];
        coords[2 * k + 1] = cxy[2 * p + 1];
        for (int a = 0; a < 4; ++a) pnodes[4 * k + a] = nd[4 * p + a];
        pelts[k] = el[p];
        i64 rr = (i64)floor((double)cxy[2 * p + 1] - 0.5), cc = (i64)floor((double)cxy[2 * p] - 0.5);
        if (rr < 0) rr = 0; if (rr > H - 1) rr = H - 1;
        if (cc < 0) cc = 0; if (cc > W - 1) cc = W - 1;
        roi_flat[k] = rr * W + cc;
        ++k;
    }
    free(cxy);
    free(el);
    free(nd);
    return k;
}

/* build_node_pixel_dense / _build_node_pixel_dense_jax (pixel_assets.py:459-526):
 * sequential scan over pixel_nodes flattened pixel-major; slot = running
 * count of the node.  With idx == NULL only degrees and max degree are
 * produced. */
int FN(orc_node_pixel_dense)(const i32* pnodes, const REAL* shapeN, i64 Np, i64 Nn, i64 maxdeg,
                             i32* degree, i64* maxdeg_out, i32* idx, REAL* wts) {
    memset(degree, 0, sizeof(i32) * (size_t)Nn);
    for (i64 t = 0; t < 4 * Np; ++t) degree[pnodes[t]]++;
    i64 md = 0;
    for (i64 k = 0; k < Nn; ++k) if (degree[k] > md) md = degree[k];
    if (maxdeg_out) *maxdeg_out = md;
    if (!idx) return 0;
    if (maxdeg < md) return -1;
    for (i64 k = 0; k < Nn * maxdeg; ++k) { idx[k] = -1; wts[k] = 0; }
    i32* cnt = (i32*)calloc((size_t)Nn, sizeof(i32));
    for (i64 t = 0; t < 4 * Np; ++t) {
        i32 nd = pnodes[t];
        i64 slot = cnt[nd]++;
        idx[(i64)nd * maxdeg + slot] = (i32)(t / 4);
        wts[(i64)nd * maxdeg + slot] = shapeN[t];
    }
    free(cnt);
    return 0;
}

/* build_node_neighbor_dense (pixel_assets.py:330-409): ordered pairs of
 * distinct element nodes, deduplicated, lexsorted by (src, dst); weight
 * 1/(|x_j - x_i| + 1e-6).  Tables are (Nn, maxdeg_cap). */
static int cmp_pair(const void* a, const void* b) {
    const i64* x = (const i64*)a;
    const i64* y = (const i64*)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
    return 0;
}

int FN(orc_node_neighbor_dense)(const i32* elems, i64 Ne, const REAL* nodes, i64 Nn, i64 maxdeg_cap,
                                i32* idx, i32* degree, REAL* wts, i64* maxdeg_out) {
    i64 m = 12 * Ne;
    i64* pairs = (i64*)malloc(sizeof(i64) * 2 * (size_t)(m + 1));
    i64 k = 0;
    for (i64 e = 0; e < Ne; ++e)
        for (int s = 0; s < 4; ++s)
            for (int d = 0; d < 4; ++d) {
                if (s == d) continue;
                pairs[2 * k] = elems[4 * e + s];
                pairs[2 * k + 1] = elems[4 * e + d];
                ++k;
            }
    qsort(pairs, (size_t)m, 2 * sizeof(i64), cmp_pair);
    memset(degree, 0, sizeof(i32) * (size_t)Nn);
    i64 md = 0;
    i64 prev_s = -1, prev_d = -1;
    for (i64 t = 0; t < m; ++t) {
        i64 s = pairs[2 * t], d = pairs[2 * t + 1];
        if (s == prev_s && d == prev_d) continue;
        prev_s = s; prev_d = d;
        if (s == d) continue;
        i64 slot = degree[s]++;
        if (degree[s] > md) md = degree[s];
        if (slot < maxdeg_cap) {
            REAL dx = nodes[2 * d] - nodes[2 * s];
            REAL dy = nodes[2 * d + 1] - nodes[2 * s + 1];
            REAL dist = (REAL)sqrt((double)(dx * dx + dy * dy));
            idx[s * maxdeg_cap + slot] = (i32)d;
            wts[s * maxdeg_cap + slot] = (REAL)1.0 / (dist + (REAL)1e-6);
        }
    }
    *maxdeg_out = md;
    free(pairs);
    return md > maxdeg_cap ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Image sampling (solver_global_cg.py:472-518)                               */
/* ------------------------------------------------------------------------ */
static inline i64 clampi(i64 v, i64 lo, i64 hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Integer cell of a floored coordinate.  Coordinates far outside the image
 * (or NaN, from a diverged iterate) would overflow the conversion; every tap
 * of such a sample is clamped to the image edge anyway, so the cell is
 * clamped to +-2^40 first (NaN: cell 0; its weights are NaN) and finite
 * in-range results are unchanged. */
static inline i64 cell_index(REAL f) {
    if (!(f == f)) return 0;
    if (f < (REAL)-1099511627776.0) return -1099511627776LL;
    if (f > (REAL)1099511627776.0) return 1099511627776LL;
    return (i64)f;
}

/* _bilinear_sample on the transposed image (solver_global_cg.py:481-509).
 * The image is (H, W) row-major.  The clip bounds come from the transposed
 * shape (x index clipped by H-1, y index by W-1) and the gather then clamps
 * to the array extent (x by W-1, y by H-1).  Returns the value and its
 * coordinate derivative (the derivative of fx, fy is 1; floor contributes 0). */
static void sample_linear(const REAL* img, i64 H, i64 W, REAL x, REAL y, REAL* v, REAL* gx, REAL* gy) {
    REAL xs = x - (REAL)0.5, ys = y - (REAL)0.5;
    REAL x0f = (REAL)floor((double)xs), y0f = (REAL)floor((double)ys);
    i64 x0 = cell_index(x0f), y0 = cell_index(y0f);
    REAL fx = xs - x0f, fy = ys - y0f;
    i64 xa = clampi(clampi(x0, 0, H - 1), 0, W - 1), xb = clampi(clampi(x0 + 1, 0, H - 1), 0, W - 1);
    i64 ya = clampi(clampi(y0, 0, W - 1), 0, H - 1), yb = clampi(clampi(y0 + 1, 0, W - 1), 0, H - 1);
    REAL I00 = img[ya * W + xa], I10 = img[ya * W + xb], I01 = img[yb * W + xa], I11 = img[yb * W + xb];
    REAL ofx = (REAL)1 - fx, ofy = (REAL)1 - fy;
    *v = ofx * ofy * I00 + fx * ofy * I10 + ofx * fy * I01 + fx * fy * I11;
    if (gx) {
        *gx = ofy * (I10 - I00) + fy * (I11 - I01);
        *gy = ofx * (I01 - I00) + fx * (I11 - I10);
    }
}

/* Cubic convolution, Keys (1981) kernel with a = -1/2 (Catmull-Rom), on
 * taps -1..2 around floor(t); boundary "nearest" (clamp-to-edge); sample
 * coordinates shifted by -0.5 as in _cubic_sample (solver_global_cg.py:512-518).
 * The reference delegates to im_jax.flat_nd_cubic_interpolate, which is not
 * available; this kernel is the build's stated choice. */
static void keys_weights(REAL t, REAL w[4], REAL dw[4]) {
    REAL t2 = t * t, t3 = t2 * t;
    w[0] = (REAL)0.5 * (-t3 + (REAL)2 * t2 - t);
    w[1] = (REAL)0.5 * ((REAL)3 * t3 - (REAL)5 * t2 + (REAL)2);
    w[2] = (REAL)0.5 * ((REAL)-3 * t3 + (REAL)4 * t2 + t);
    w[3] = (REAL)0.5 * (t3 - t2);
    if (dw) {
        dw[0] = (REAL)0.5 * ((REAL)-3 * t2 + (REAL)4 * t - (REAL)1);
        dw[1] = (REAL)0.5 * ((REAL)9 * t2 - (REAL)10 * t);
        dw[2] = (REAL)0.5 * ((REAL)-9 * t2 + (REAL)8 * t + (REAL)1);
        dw[3] = (REAL)0.5 * ((REAL)3 * t2 - (REAL)2 * t);
    }
}

static void sample_cubic(const REAL* img, i64 H, i64 W, REAL x, REAL y, REAL* v, REAL* gx, REAL* gy) {
    REAL xs = x - (REAL)0.5, ys = y - (REAL)0.5;
    REAL x0f = (REAL)floor((double)xs), y0f = (REAL)floor((double)ys);
    i64 x0 = cell_index(x0f), y0 = cell_index(y0f);
    REAL tx = xs - x0f, ty = ys - y0f;
    REAL wx[4], dwx[4], wy[4], dwy[4];
    keys_weights(tx, wx, dwx);
    keys_weights(ty, wy, dwy);
    REAL acc = 0, accx = 0, accy = 0;
    for (int i = 0; i < 4; ++i) {
        i64 yy = clampi(y0 - 1 + i, 0, H - 1);
        REAL rv = 0, rd = 0;
        for (int j = 0; j < 4; ++j) {
            i64 xx = clampi(x0 - 1 + j, 0, W - 1);
            REAL I = img[yy * W + xx];
            rv = rv + wx[j] * I;
            rd = rd + dwx[j] * I;
        }
        acc = acc + wy[i] * rv;
        accx = accx + wy[i] * rd;
        accy = accy + dwy[i] * rv;
    }
    *v = acc;
    if (gx) {
        *gx = accx;
        *gy = accy;
    }
}

static inline void sample(int interp, const REAL* img, i64 H, i64 W, REAL x, REAL y, REAL* v, REAL* gx, REAL* gy) {
    if (interp == 0) sample_cubic(img, H, W, x, y, v, gx, gy);
    else sample_linear(img, H, W, x, y, v, gx, gy);
}

/* interp: 0 = cubic, 1 = linear.  gx/gy may be NULL. */
int FN(orc_sample)(const REAL* img, i64 H, i64 W, const REAL* xy, i64 n, int interp, REAL* v, REAL* gx, REAL* gy) {
#pragma omp parallel for schedule(static)
    for (i64 p = 0; p < n; ++p) {
        REAL a, b, c;
        sample(interp, img, H, W, xy[2 * p], xy[2 * p + 1], &a, &b, &c);
        v[p] = a;
        if (gx) { gx[p] = b; gy[p] = c; }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Objective (solver_global_cg.py:205-323)                                    */
/* ------------------------------------------------------------------------ */
typedef struct {
    const REAL* coords;   /* (Np,2) */
    const i32* pnodes;    /* (Np,4) */
    const REAL* shapeN;   /* (Np,4) */
    const i32* npix_idx;  /* (Nn,maxdeg_pix) node->pixel table */
    const REAL* npix_w;
    const i32* npix_deg;
    i64 maxdeg_pix;
    const i32* nb_idx;    /* (Nn,maxdeg_nb) regularization graph */
    const i32* nb_deg;
    const REAL* nb_w;
    const REAL* reg_w;    /* (Nn,) */
    i64 maxdeg_nb;
    i64 Np, Nn;
} assets_t;

/* reg_energy_spring_global (solver_global_cg.py:261-286) and its exact
 * gradient (both halves of each ordered pair). */
static double reg_energy(const assets_t* A, const REAL* u, REAL* grad_acc /* may be NULL; adds alpha*dE */, REAL alpha) {
    double E = 0;
    for (i64 i = 0; i < A->Nn; ++i) {
        i64 deg = A->nb_deg[i];
        double Ei = 0;
        for (i64 k = 0; k < A->maxdeg_nb; ++k) {
            if (k >= deg) break;
            i64 j = A->nb_idx[i * A->maxdeg_nb + k];
            REAL w = A->nb_w[i * A->maxdeg_nb + k] * (REAL)0.5 * (A->reg_w[i] + A->reg_w[j]);
            REAL dx = u[2 * i] - u[2 * j], dy = u[2 * i + 1] - u[2 * j + 1];
            Ei += (double)(w * (dx * dx + dy * dy));
        }
        E += Ei;
    }
    if (grad_acc) {
        for (i64 i = 0; i < A->Nn; ++i) {
            i64 deg = A->nb_deg[i];
            for (i64 k = 0; k < deg && k < A->maxdeg_nb; ++k) {
                i64 j = A->nb_idx[i * A->maxdeg_nb + k];
                REAL w = A->nb_w[i * A->maxdeg_nb + k] * (REAL)0.5 * (A->reg_w[i] + A->reg_w[j]);
                REAL tx = alpha * (REAL)0.5 * w * (u[2 * i] - u[2 * j]);
                REAL ty = alpha * (REAL)0.5 * w * (u[2 * i + 1] - u[2 * j + 1]);
                grad_acc[2 * i] += tx; grad_acc[2 * i + 1] += ty;
                grad_acc[2 * j] -= tx; grad_acc[2 * j + 1] -= ty;
            }
        }
    }
    return 0.25 * E;
}

#define CHUNK 4096

/* J_total and (optionally) its gradient (solver_global_cg.py:289-323).
 * fref: precomputed f(x_ref) per pixel.  q: scratch (Np,2).  The image
 * gradient part is gathered through the node->pixel table, so the result is
 * independent of the thread count. */
static double objective(const assets_t* A, const REAL* u, const REAL* fref, const REAL* gimg, i64 H, i64 W,
                        int interp, REAL alpha, REAL* grad, REAL* q) {
    i64 nch = (A->Np + CHUNK - 1) / CHUNK;
    double* part = (double*)calloc((size_t)(nch + 1), sizeof(double));
    REAL invN = (REAL)1 / (REAL)(A->Np > 0 ? A->Np : 1);
#pragma omp parallel for schedule(static)
    for (i64 c = 0; c < nch; ++c) {
        double s = 0;
        i64 p1 = (c + 1) * CHUNK < A->Np ? (c + 1) * CHUNK : A->Np;
        for (i64 p = c * CHUNK; p < p1; ++p) {
            REAL ux = 0, uy = 0;
            for (int a = 0; a < 4; ++a) {
                i64 nd = A->pnodes[4 * p + a];
                ux = ux + A->shapeN[4 * p + a] * u[2 * nd];
                uy = uy + A->shapeN[4 * p + a] * u[2 * nd + 1];
            }
            REAL xd = A->coords[2 * p] + ux, yd = A->coords[2 * p + 1] + uy;
            REAL g, gx, gy;
            sample(interp, gimg, H, W, xd, yd, &g, grad ? &gx : NULL, grad ? &gy : NULL);
            REAL r = g - fref[p];
            s += (double)r * (double)r;
            if (grad) {
                q[2 * p] = r * gx * invN;
                q[2 * p + 1] = r * gy * invN;
            }
        }
        part[c] = s;
    }
    double S = 0;
    for (i64 c = 0; c < nch; ++c) S += part[c];
    free(part);
    double Jimg = 0.5 * S / (double)(A->Np > 0 ? A->Np : 1);
    if (grad) {
#pragma omp parallel for schedule(static)
        for (i64 n = 0; n < A->Nn; ++n) {
            double gxs = 0, gys = 0;
            i64 deg = A->npix_deg[n];
            for (i64 k = 0; k < deg; ++k) {
                i64 p = A->npix_idx[n * A->maxdeg_pix + k];
                REAL w = A->npix_w[n * A->maxdeg_pix + k];
                gxs += (double)(w * q[2 * p]);
                gys += (double)(w * q[2 * p + 1]);
            }
            grad[2 * n] = (REAL)gxs;
            grad[2 * n + 1] = (REAL)gys;
        }
    }
    REAL Er = (REAL)reg_energy(A, u, grad, alpha);
    REAL Jr = (REAL)Jimg;
    return (double)(Jr + alpha * Er);
}

static assets_t make_assets(const REAL* coords, const i32* pnodes, const REAL* shapeN, i64 Np,
                            const i32* npix_idx, const REAL* npix_w, const i32* npix_deg, i64 maxdeg_pix,
                            const i32* nb_idx, const i32* nb_deg, const REAL* nb_w, const REAL* reg_w,
                            i64 maxdeg_nb, i64 Nn) {
    assets_t A;
    A.coords = coords; A.pnodes = pnodes; A.shapeN = shapeN; A.Np = Np;
    A.npix_idx = npix_idx; A.npix_w = npix_w; A.npix_deg = npix_deg; A.maxdeg_pix = maxdeg_pix;
    A.nb_idx = nb_idx; A.nb_deg = nb_deg; A.nb_w = nb_w; A.reg_w = reg_w; A.maxdeg_nb = maxdeg_nb;
    A.Nn = Nn;
    return A;
}

/* J_total / value_and_grad at u (grad may be NULL). */
double FN(orc_objective)(const REAL* u, const REAL* fimg, const REAL* gimg, i64 H, i64 W, int interp, REAL alpha,
                         const REAL* coords, const i32* pnodes, const REAL* shapeN, i64 Np,
                         const i32* npix_idx, const REAL* npix_w, const i32* npix_deg, i64 maxdeg_pix,
                         const i32* nb_idx, const i32* nb_deg, const REAL* nb_w, const REAL* reg_w, i64 maxdeg_nb,
                         i64 Nn, REAL* grad) {
    assets_t A = make_assets(coords, pnodes, shapeN, Np, npix_idx, npix_w, npix_deg, maxdeg_pix, nb_idx, nb_deg,
                             nb_w, reg_w, maxdeg_nb, Nn);
    REAL* fref = (REAL*)malloc(sizeof(REAL) * (size_t)(Np + 1));
    FN(orc_sample)(fimg, H, W, coords, Np, interp, fref, NULL, NULL);
    REAL* q = grad ? (REAL*)malloc(sizeof(REAL) * 2 * (size_t)(Np + 1)) : NULL;
    double J = objective(&A, u, fref, gimg, H, W, interp, alpha, grad, q);
    free(fref);
    free(q);
    return J;
}

/* _cg_solve (solver_global_cg.py:326-469): NLCG with Polak-Ribiere-plus beta,
 * restart on non-descent, 10-trial Armijo backtracking (c = 1e-4) that falls
 * through with alpha = 2^-10, stop when |grad|_2 < tol before the update.
 * hist: (max_iter, 3) rows [J_trial, |grad|, alpha], NaN padded. */
int FN(orc_cg_solve)(REAL* u, const REAL* fimg, const REAL* gimg, i64 H, i64 W, int interp, REAL alpha_reg,
                     int max_iter, REAL tol,
                     const REAL* coords, const i32* pnodes, const REAL* shapeN, i64 Np,
                     const i32* npix_idx, const REAL* npix_w, const i32* npix_deg, i64 maxdeg_pix,
                     const i32* nb_idx, const i32* nb_deg, const REAL* nb_w, const REAL* reg_w, i64 maxdeg_nb,
                     i64 Nn, REAL* hist, int* n_iters) {
    assets_t A = make_assets(coords, pnodes, shapeN, Np, npix_idx, npix_w, npix_deg, maxdeg_pix, nb_idx, nb_deg,
                             nb_w, reg_w, maxdeg_nb, Nn);
    i64 n2 = 2 * Nn;
    REAL* fref = (REAL*)malloc(sizeof(REAL) * (size_t)(Np + 1));
    FN(orc_sample)(fimg, H, W, coords, Np, interp, fref, NULL, NULL);
    REAL* q = (REAL*)malloc(sizeof(REAL) * 2 * (size_t)(Np + 1));
    REAL* g = (REAL*)calloc((size_t)n2, sizeof(REAL));
    REAL* gp = (REAL*)calloc((size_t)n2, sizeof(REAL));
    REAL* d = (REAL*)calloc((size_t)n2, sizeof(REAL));
    REAL* dn = (REAL*)calloc((size_t)n2, sizeof(REAL));
    REAL* cand = (REAL*)calloc((size_t)n2, sizeof(REAL));
    for (int k = 0; k < max_iter; ++k) {
        hist[3 * k] = hist[3 * k + 1] = hist[3 * k + 2] = (REAL)NAN;
    }
    int k = 0;
    while (k < max_iter) {
        REAL Jv = (REAL)objective(&A, u, fref, gimg, H, W, interp, alpha_reg, g, q);
        double gg = 0;
        for (i64 t = 0; t < n2; ++t) gg += (double)g[t] * (double)g[t];
        REAL gnorm = (REAL)sqrt(gg);
        if (gnorm < tol) {
            hist[3 * k] = Jv; hist[3 * k + 1] = gnorm; hist[3 * k + 2] = 0;
            break;
        }
        if (k == 0) {
            for (i64 t = 0; t < n2; ++t) dn[t] = -g[t];
        } else {
            double num = 0, den = 0;
            for (i64 t = 0; t < n2; ++t) {
                num += (double)(g[t] * (g[t] - gp[t]));
                den += (double)gp[t] * (double)gp[t];
            }
            REAL beta = (REAL)num / ((REAL)den + (REAL)1e-16);
            if (beta < 0) beta = 0; /* jnp.maximum keeps NaN */
            for (i64 t = 0; t < n2; ++t) dn[t] = -g[t] + beta * d[t];
        }
        double gd = 0;
        for (i64 t = 0; t < n2; ++t) gd += (double)g[t] * (double)dn[t];
        REAL gdr = (REAL)gd;
        if (gdr >= 0) {
            for (i64 t = 0; t < n2; ++t) dn[t] = -g[t];
            gdr = (REAL)(-gg);
        }
        REAL alpha = 1, Jt = Jv;
        int done = 0;
        for (int ls = 0; ls < 10; ++ls) {
            if (done) break;
            for (i64 t = 0; t < n2; ++t) cand[t] = u[t] + alpha * dn[t];
            Jt = (REAL)objective(&A, cand, fref, gimg, H, W, interp, alpha_reg, NULL, NULL);
            int ok = Jt <= Jv + (REAL)1e-4 * alpha * gdr;
            if (ok) done = 1;
            else alpha = alpha * (REAL)0.5;
        }
        for (i64 t = 0; t < n2; ++t) u[t] = u[t] + alpha * dn[t];
        hist[3 * k] = Jt; hist[3 * k + 1] = gnorm; hist[3 * k + 2] = alpha;
        REAL* tmp = d; d = dn; dn = tmp;
        tmp = gp; gp = g; g = tmp;
        ++k;
    }
    *n_iters = k;
    free(fref); free(q); free(g); free(gp); free(d); free(dn); free(cand);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Translation ZNCC (solver_translation_zncc.py:89-168)                        */
/* ------------------------------------------------------------------------ */
static REAL zncc_one(const REAL* t_hat, const REAL* img, i64 W, i64 r0, i64 c0, int win) {
    i64 n = (i64)win * win;
    double s = 0;
    for (int i = 0; i < win; ++i)
        for (int j = 0; j < win; ++j) s += img[(r0 + i) * W + c0 + j];
    REAL mean = (REAL)(s / (double)n);
    double v = 0;
    for (int i = 0; i < win; ++i)
        for (int j = 0; j < win; ++j) {
            REAL dlt = img[(r0 + i) * W + c0 + j] - mean;
            v += (double)(dlt * dlt);
        }
    REAL sd = (REAL)sqrt(v / (double)n);
    REAL den = sd + (REAL)1e-12;
    double acc = 0;
    for (int i = 0; i < win; ++i)
        for (int j = 0; j < win; ++j) {
            REAL wv = (img[(r0 + i) * W + c0 + j] - mean) / den;
            acc += (double)(wv * t_hat[i * win + j]);
        }
    return (REAL)(acc / (double)n);
}

int FN(orc_zncc)(const REAL* centers, i64 Nc, const REAL* ref, i64 Hr, i64 Wr, const REAL* def, i64 Hd, i64 Wd,
                 int win, int search, REAL* u_out, REAL* score_out) {
    int r = win / 2;
    int P = win + 2 * search;
    int nwin = 2 * search + 1;
#pragma omp parallel for schedule(dynamic, 4)
    for (i64 c = 0; c < Nc; ++c) {
        i64 x = (i64)nearbyint((double)centers[2 * c]);
        i64 y = (i64)nearbyint((double)centers[2 * c + 1]);
        int tpl_ok = (y - r >= 0) && (y + r < Hr) && (x - r >= 0) && (x + r < Wr);
        i64 xs = x - search, ys = y - search, xe = x + search, ye = y + search;
        int patch_ok = (ys - r >= 0) && (ye + r < Hd) && (xs - r >= 0) && (xe + r < Wd);
        if (!(tpl_ok && patch_ok)) {
            u_out[2 * c] = 0;
            u_out[2 * c + 1] = 0;
            score_out[c] = (REAL)NAN;
            continue;
        }
        i64 tr = clampi(y - r, 0, Hr - win), tc = clampi(x - r, 0, Wr - win);
        i64 pr = clampi(ys - r, 0, Hd - P), pc = clampi(xs - r, 0, Wd - P);
        REAL* t_hat = (REAL*)malloc(sizeof(REAL) * (size_t)win * win);
        i64 n = (i64)win * win;
        double s = 0;
        for (int i = 0; i < win; ++i)
            for (int j = 0; j < win; ++j) s += ref[(tr + i) * Wr + tc + j];
        REAL mean = (REAL)(s / (double)n);
        double v = 0;
        for (int i = 0; i < win; ++i)
            for (int j = 0; j < win; ++j) {
                REAL dlt = ref[(tr + i) * Wr + tc + j] - mean;
                v += (double)(dlt * dlt);
            }
        REAL den = (REAL)sqrt(v / (double)n) + (REAL)1e-12;
        for (int i = 0; i < win; ++i)
            for (int j = 0; j < win; ++j) t_hat[i * win + j] = (ref[(tr + i) * Wr + tc + j] - mean) / den;
        /* jnp.argmax: first maximum, a NaN counts as the maximum */
        REAL best = 0;
        int bi = -1;
        for (int ky = 0; ky < nwin; ++ky)
            for (int kx = 0; kx < nwin; ++kx) {
                REAL sc = zncc_one(t_hat, def, Wd, pr + ky, pc + kx, win);
                if (bi < 0) { best = sc; bi = 0; continue; }
                if (isnan(best)) continue;
                if (isnan(sc) || sc > best) { best = sc; bi = ky * nwin + kx; }
            }
        free(t_hat);
        int ky = bi / nwin, kx = bi % nwin;
        u_out[2 * c] = (REAL)((xs + kx) - x);
        u_out[2 * c + 1] = (REAL)((ys + ky) - y);
        score_out[c] = best;
    }
    return 0;
}

/* _project_centers_to_nodes_idw (dic_init_motion.py:64-91): global inverse
 * distance weighting with power 2; a centre within 1e-6 of a node is copied. */
int FN(orc_idw)(const REAL* nodes, i64 Nn, const REAL* centers, const REAL* disp, i64 Nc, REAL* out) {
#pragma omp parallel for schedule(static)
    for (i64 n = 0; n < Nn; ++n) {
        REAL nx = nodes[2 * n], ny = nodes[2 * n + 1];
        i64 close = -1;
        for (i64 c = 0; c < Nc; ++c) {
            REAL dx = centers[2 * c] - nx, dy = centers[2 * c + 1] - ny;
            REAL dist = (REAL)sqrt((double)(dx * dx + dy * dy));
            if (dist < (REAL)1e-6) { close = c; break; }
        }
        if (close >= 0) {
            out[2 * n] = disp[2 * close];
            out[2 * n + 1] = disp[2 * close + 1];
            continue;
        }
        double ws = 0, ax = 0, ay = 0;
        for (i64 c = 0; c < Nc; ++c) {
            REAL dx = centers[2 * c] - nx, dy = centers[2 * c + 1] - ny;
            REAL dist = (REAL)sqrt((double)(dx * dx + dy * dy));
            REAL dd = dist + (REAL)1e-12;
            REAL w = (REAL)1 / (dd * dd);
            ws += w;
            ax += (double)w * disp[2 * c];
            ay += (double)w * disp[2 * c + 1];
        }
        out[2 * n] = (REAL)(ax / ws);
        out[2 * n + 1] = (REAL)(ay / ws);
    }
    return 0;
}

/* compute_green_lagrange_strain_nodes_lsq + green_lagrange_to_voigt
 * (strain.py:11-114).  F_out, E_out: (Nn,2,2); voigt: (Nn,3). */
int FN(orc_strain)(const REAL* u, const REAL* X, i64 Nn, const i32* nb_idx, const i32* nb_deg, i64 maxdeg,
                   REAL L, REAL eps, REAL* F_out, REAL* E_out, REAL* voigt) {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < Nn; ++i) {
        REAL a00 = 0, a01 = 0, a10 = 0, a11 = 0, b0x = 0, b0y = 0, b1x = 0, b1y = 0;
        i64 deg = nb_deg[i];
        for (i64 k = 0; k < maxdeg; ++k) {
            int valid = k < deg;
            i64 j = valid ? nb_idx[i * maxdeg + k] : 0;
            REAL dX = X[2 * j] - X[2 * i], dY = X[2 * j + 1] - X[2 * i + 1];
            REAL dux = u[2 * j] - u[2 * i], duy = u[2 * j + 1] - u[2 * i + 1];
            REAL w = valid ? (REAL)1 : (REAL)0;
            if (L > 0) {
                REAL rr = (REAL)sqrt((double)(dX * dX + dY * dY));
                REAL z = rr / (L + (REAL)1e-12);
                w = w * (REAL)exp(-(double)(z * z));
            }
            REAL wx = dX * w, wy = dY * w;
            a00 = a00 + wx * dX; a01 = a01 + wx * dY; a10 = a10 + wy * dX; a11 = a11 + wy * dY;
            b0x = b0x + wx * dux; b0y = b0y + wy * dux;
            b1x = b1x + wx * duy; b1y = b1y + wy * duy;
        }
        a00 = a00 + eps; a11 = a11 + eps;
        REAL gxx, gxy, gyx, gyy;
        solve2(a00, a01, a10, a11, b0x, b0y, &gxx, &gxy);
        solve2(a00, a01, a10, a11, b1x, b1y, &gyx, &gyy);
        REAL F00 = (REAL)1 + gxx, F01 = gxy, F10 = gyx, F11 = (REAL)1 + gyy;
        REAL C00 = F00 * F00 + F10 * F10, C01 = F00 * F01 + F10 * F11;
        REAL C10 = F01 * F00 + F11 * F10, C11 = F01 * F01 + F11 * F11;
        REAL E00 = (REAL)0.5 * (C00 - 1), E01 = (REAL)0.5 * C01, E10 = (REAL)0.5 * C10, E11 = (REAL)0.5 * (C11 - 1);
        if (F_out) { F_out[4 * i] = F00; F_out[4 * i + 1] = F01; F_out[4 * i + 2] = F10; F_out[4 * i + 3] = F11; }
        if (E_out) { E_out[4 * i] = E00; E_out[4 * i + 1] = E01; E_out[4 * i + 2] = E10; E_out[4 * i + 3] = E11; }
        if (voigt) {
            voigt[3 * i] = E00;
            voigt[3 * i + 1] = E11;
            voigt[3 * i + 2] = (REAL)0.5 * (E01 + E10);
        }
    }
    return 0;
}

int FN(orc_num_threads)(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
