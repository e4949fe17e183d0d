// Image samplers shared by the solver kernels.
//
// Coordinates are pixel-centre based: pixel (i, j) has centre (j + 0.5, i + 0.5),
// so both samplers shift by -0.5 into index space (solver_global_cg.py:485-488,
// 515-517).  Images are (H, W) float32, row-major.
#pragma once

#include <hip/hip_runtime.h>

namespace d2ic {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Samplers take the sample position as an integer cell index plus a
// fractional offset in [0, 1): position - 0.5 = x0 + fx.  Callers that know the
// reference coordinate of a pixel split it as integer part + small offset so
// the fractional part keeps full FP32 precision even where |x| ~ 4096 (an
// absolute FP32 coordinate there is quantised to ~2.4e-4 px).

// Bilinear sampler of _bilinear_sample (solver_global_cg.py:481-509) with its
// clip bounds: the x index is clipped by H-1 and the y index by W-1 (the bounds
// come from the transposed image), then the gather clamps x by W-1 and y by
// H-1.  Returns value and coordinate derivative (floor contributes nothing).
__device__ __forceinline__ void sample_linear_at(const float* __restrict__ img_, int H, int W, int x0, float fx,
                                                 int y0, float fy, float& v, float& gx, float& gy) {
    const __attribute__((address_space(1))) float* img = (const __attribute__((address_space(1))) float*)img_;
    const int xa = clampi(clampi(x0, 0, H - 1), 0, W - 1), xb = clampi(clampi(x0 + 1, 0, H - 1), 0, W - 1);
    const int ya = clampi(clampi(y0, 0, W - 1), 0, H - 1), yb = clampi(clampi(y0 + 1, 0, W - 1), 0, H - 1);
    const float I00 = img[(size_t)ya * W + xa], I10 = img[(size_t)ya * W + xb];
    const float I01 = img[(size_t)yb * W + xa], I11 = img[(size_t)yb * W + xb];
    const float ofx = 1.0f - fx, ofy = 1.0f - fy;
    v = ofx * ofy * I00 + fx * ofy * I10 + ofx * fy * I01 + fx * fy * I11;
    gx = ofy * (I10 - I00) + fy * (I11 - I01);
    gy = ofx * (I01 - I00) + fx * (I11 - I10);
}

// Cubic convolution with the Keys (1981) kernel, a = -1/2 (Catmull-Rom):
// weights of taps floor(t)-1 .. floor(t)+2 for fractional offset t and their
// t-derivatives.  Boundary "nearest": tap indices clamp to the image.
__device__ __forceinline__ void keys_weights(float t, float w[4], float dw[4]) {
    // Horner forms of w0 = (-t^3 + 2t^2 - t)/2, w1 = (3t^3 - 5t^2 + 2)/2,
    // w2 = (-3t^3 + 4t^2 + t)/2, w3 = (t^3 - t^2)/2 and their derivatives
    const float t2 = t * t, h = 0.5f * t;
    const float p2 = __fmaf_rn(-1.5f, t, 2.0f);  // shared by w2 and dw0
    w[0] = __fmaf_rn(__fmaf_rn(-0.5f, t, 1.0f), t, -0.5f) * t;
    w[1] = __fmaf_rn(__fmaf_rn(1.5f, t, -2.5f), t2, 1.0f);
    w[2] = __fmaf_rn(p2, t2, h);
    w[3] = __fmaf_rn(0.5f, t, -0.5f) * t2;
    dw[0] = __fmaf_rn(p2, t, -0.5f);
    dw[1] = __fmaf_rn(4.5f, t, -5.0f) * t;
    dw[2] = __fmaf_rn(__fmaf_rn(-4.5f, t, 4.0f), t, 0.5f);
    dw[3] = __fmaf_rn(1.5f, t, -1.0f) * t;
}

__device__ __forceinline__ void sample_cubic_at(const float* __restrict__ img_, int H, int W, int x0, float tx, int y0,
                                                float ty, float& v, float& gx, float& gy) {
    const __attribute__((address_space(1))) float* img = (const __attribute__((address_space(1))) float*)img_;
    float wx[4], dwx[4], wy[4], dwy[4];
    keys_weights(tx, wx, dwx);
    keys_weights(ty, wy, dwy);
    int cx[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) cx[j] = clampi(x0 - 1 + j, 0, W - 1);
    float acc = 0.f, accx = 0.f, accy = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const __attribute__((address_space(1))) float* row = img + (size_t)clampi(y0 - 1 + i, 0, H - 1) * W;
        float rv = 0.f, rd = 0.f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float I = row[cx[j]];
            rv += wx[j] * I;
            rd += dwx[j] * I;
        }
        acc += wy[i] * rv;
        accx += wy[i] * rd;
        accy += dwy[i] * rv;
    }
    v = acc;
    gx = accx;
    gy = accy;
}

__device__ __forceinline__ void sample_at(int interp, const float* __restrict__ img, int H, int W, int x0, float fx,
                                          int y0, float fy, float& v, float& gx, float& gy) {
    if (interp == 0)
        sample_cubic_at(img, H, W, x0, fx, y0, fy, v, gx, gy);
    else
        sample_linear_at(img, H, W, x0, fx, y0, fy, v, gx, gy);
}

// Tap index clamps of the two samplers (see sample_linear_at / sample_cubic_at).
__device__ __forceinline__ int clamp_col(int interp, int x, int H, int W) {
    return interp == 0 ? clampi(x, 0, W - 1) : clampi(clampi(x, 0, H - 1), 0, W - 1);
}
__device__ __forceinline__ int clamp_row(int interp, int y, int H, int W) {
    return interp == 0 ? clampi(y, 0, H - 1) : clampi(clampi(y, 0, W - 1), 0, H - 1);
}

// Same samplers reading a staged box: box[(r) * bw + c] holds
// img[clamp_row(by0 + r), clamp_col(bx0 + c)]; (lx, ly) = (x0 - bx0, y0 - by0).
__device__ __forceinline__ void sample_box(int interp, const __attribute__((address_space(3))) float* box, int bw, int lx,
                                           float fx, int ly, float fy, float& v, float& gx, float& gy) {
    if (interp == 0) {
        // all 16 taps are requested before the weights are formed, so the LDS
        // latency is paid once per sample rather than once per tap row
        float I4[4][4];
        const __attribute__((address_space(3))) float* base = box + (ly - 1) * bw + (lx - 1);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) I4[i][j] = base[i * bw + j];
        float wx[4], dwx[4], wy[4], dwy[4];
        keys_weights(fx, wx, dwx);
        keys_weights(fy, wy, dwy);
        float acc = 0.f, accx = 0.f, accy = 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float rv = 0.f, rd = 0.f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float I = I4[i][j];
                rv += wx[j] * I;
                rd += dwx[j] * I;
            }
            acc += wy[i] * rv;
            accx += wy[i] * rd;
            accy += dwy[i] * rv;
        }
        v = acc;
        gx = accx;
        gy = accy;
    } else {
        const __attribute__((address_space(3))) float* r0 = box + ly * bw + lx;
        const float I00 = r0[0], I10 = r0[1], I01 = r0[bw], I11 = r0[bw + 1];
        const float ofx = 1.0f - fx, ofy = 1.0f - fy;
        v = ofx * ofy * I00 + fx * ofy * I10 + ofx * fy * I01 + fx * fy * I11;
        gx = ofy * (I10 - I00) + fy * (I11 - I01);
        gy = ofx * (I01 - I00) + fx * (I11 - I10);
    }
}

// Split of a pixel coordinate plus displacement: (x - 0.5) + u = x0 + f with
// x = xi + xf (xi integer), evaluated as xi + ((xf - 0.5) + u).
__device__ __forceinline__ void split_coord(float x, float u, int& x0, float& f) {
    const float xi = floorf(x);
    const float t = (x - xi - 0.5f) + u;
    const float ti = floorf(t);
    x0 = (int)xi + (int)ti;
    f = t - ti;
}

// Second-derivative weights of the Keys a = -1/2 kernel (taps floor(t)-1 ..
// floor(t)+2), for the Hessian of the interpolant used by the adjoint.
__device__ __forceinline__ void keys_weights2(float t, float d2w[4]) {
    d2w[0] = 0.5f * (-6.0f * t + 4.0f);
    d2w[1] = 0.5f * (18.0f * t - 10.0f);
    d2w[2] = 0.5f * (-18.0f * t + 8.0f);
    d2w[3] = 0.5f * (6.0f * t - 2.0f);
}

// Taps of either sampler at cell (x0, y0) with fractional offsets (fx, fy):
// up to 4 x 4 clamped pixel indices with weights w and weight gradients
// (dw/dx, dw/dy) such that value = sum w * I and grad = sum (dw/dx, dw/dy) * I.
// Returns the number of taps per axis (4 cubic, 2 linear).
__device__ __forceinline__ int sampler_taps(int interp, int H, int W, int x0, float fx, int y0, float fy, int col[4],
                                            int row[4], float wx[4], float dwx[4], float wy[4], float dwy[4]) {
    if (interp == 0) {
        keys_weights(fx, wx, dwx);
        keys_weights(fy, wy, dwy);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            col[j] = clampi(x0 - 1 + j, 0, W - 1);
            row[j] = clampi(y0 - 1 + j, 0, H - 1);
        }
        return 4;
    }
    wx[0] = 1.0f - fx;
    wx[1] = fx;
    dwx[0] = -1.0f;
    dwx[1] = 1.0f;
    wy[0] = 1.0f - fy;
    wy[1] = fy;
    dwy[0] = -1.0f;
    dwy[1] = 1.0f;
    col[0] = clamp_col(1, x0, H, W);
    col[1] = clamp_col(1, x0 + 1, H, W);
    row[0] = clamp_row(1, y0, H, W);
    row[1] = clamp_row(1, y0 + 1, H, W);
    return 2;
}

// Value, gradient and Hessian (hxx, hxy, hyy) of the interpolant.
__device__ __forceinline__ void sample_hess(int interp, const float* __restrict__ img, int H, int W, int x0, float fx,
                                            int y0, float fy, float out[6]) {
    int col[4], row[4];
    float wx[4], dwx[4], wy[4], dwy[4], d2x[4] = {0.f, 0.f, 0.f, 0.f}, d2y[4] = {0.f, 0.f, 0.f, 0.f};
    const int n = sampler_taps(interp, H, W, x0, fx, y0, fy, col, row, wx, dwx, wy, dwy);
    if (n == 4) {
        keys_weights2(fx, d2x);
        keys_weights2(fy, d2y);
    }
    float v = 0.f, gx = 0.f, gy = 0.f, hxx = 0.f, hxy = 0.f, hyy = 0.f;
    for (int i = 0; i < n; ++i) {
        float rv = 0.f, rd = 0.f, rdd = 0.f;
        for (int j = 0; j < n; ++j) {
            const float I = img[(size_t)row[i] * W + col[j]];
            rv += wx[j] * I;
            rd += dwx[j] * I;
            rdd += d2x[j] * I;
        }
        v += wy[i] * rv;
        gx += wy[i] * rd;
        gy += dwy[i] * rv;
        hxx += wy[i] * rdd;
        hxy += dwy[i] * rd;
        hyy += d2y[i] * rv;
    }
    out[0] = v;
    out[1] = gx;
    out[2] = gy;
    out[3] = hxx;
    out[4] = hxy;
    out[5] = hyy;
}

// Sample at an absolute coordinate (reference form: floor(x - 0.5)).
__device__ __forceinline__ void sample(int interp, const float* __restrict__ img, int H, int W, float x, float y,
                                       float& v, float& gx, float& gy) {
    int x0, y0;
    float fx, fy;
    split_coord(x, 0.f, x0, fx);
    split_coord(y, 0.f, y0, fy);
    sample_at(interp, img, H, W, x0, fx, y0, fy, v, gx, gy);
}

}  // namespace d2ic
