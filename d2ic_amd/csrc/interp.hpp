// Image samplers shared by the solver kernels.
//
// Coordinates are pixel-centre based: pixel (i, j) has centre (j + 0.5, i + 0.5),
// so both samplers shift by -0.5 into index space (solver_global_cg.py:485-488,
// 515-517).  Images are (H, W) float32, row-major.
#pragma once

#include <hip/hip_runtime.h>

namespace d2ic {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Bilinear sampler of _bilinear_sample (solver_global_cg.py:481-509) with its
// clip bounds: the x index is clipped by H-1 and the y index by W-1 (the bounds
// come from the transposed image), then the gather clamps x by W-1 and y by
// H-1.  Returns value and coordinate derivative (floor contributes nothing).
__device__ __forceinline__ void sample_linear(const float* __restrict__ img, int H, int W, float x, float y,
                                              float& v, float& gx, float& gy) {
    const float xs = x - 0.5f, ys = y - 0.5f;
    const float x0f = floorf(xs), y0f = floorf(ys);
    const int x0 = (int)x0f, y0 = (int)y0f;
    const float fx = xs - x0f, fy = ys - y0f;
    const int xa = clampi(clampi(x0, 0, H - 1), 0, W - 1), xb = clampi(clampi(x0 + 1, 0, H - 1), 0, W - 1);
    const int ya = clampi(clampi(y0, 0, W - 1), 0, H - 1), yb = clampi(clampi(y0 + 1, 0, W - 1), 0, H - 1);
    const float I00 = img[ya * W + xa], I10 = img[ya * W + xb];
    const float I01 = img[yb * W + xa], I11 = img[yb * W + xb];
    const float ofx = 1.0f - fx, ofy = 1.0f - fy;
    v = ofx * ofy * I00 + fx * ofy * I10 + ofx * fy * I01 + fx * fy * I11;
    gx = ofy * (I10 - I00) + fy * (I11 - I01);
    gy = ofx * (I01 - I00) + fx * (I11 - I10);
}

// Cubic convolution with the Keys (1981) kernel, a = -1/2 (Catmull-Rom):
// weights of taps floor(t)-1 .. floor(t)+2 for fractional offset t and their
// t-derivatives.  Boundary "nearest": tap indices clamp to the image.
__device__ __forceinline__ void keys_weights(float t, float w[4], float dw[4]) {
    const float t2 = t * t, t3 = t2 * t;
    w[0] = 0.5f * (-t3 + 2.0f * t2 - t);
    w[1] = 0.5f * (3.0f * t3 - 5.0f * t2 + 2.0f);
    w[2] = 0.5f * (-3.0f * t3 + 4.0f * t2 + t);
    w[3] = 0.5f * (t3 - t2);
    dw[0] = 0.5f * (-3.0f * t2 + 4.0f * t - 1.0f);
    dw[1] = 0.5f * (9.0f * t2 - 10.0f * t);
    dw[2] = 0.5f * (-9.0f * t2 + 8.0f * t + 1.0f);
    dw[3] = 0.5f * (3.0f * t2 - 2.0f * t);
}

__device__ __forceinline__ void sample_cubic(const float* __restrict__ img, int H, int W, float x, float y, float& v,
                                             float& gx, float& gy) {
    const float xs = x - 0.5f, ys = y - 0.5f;
    const float x0f = floorf(xs), y0f = floorf(ys);
    const int x0 = (int)x0f, y0 = (int)y0f;
    float wx[4], dwx[4], wy[4], dwy[4];
    keys_weights(xs - x0f, wx, dwx);
    keys_weights(ys - y0f, wy, dwy);
    int cx[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) cx[j] = clampi(x0 - 1 + j, 0, W - 1);
    float acc = 0.f, accx = 0.f, accy = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float* row = img + (size_t)clampi(y0 - 1 + i, 0, H - 1) * W;
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

__device__ __forceinline__ void sample(int interp, const float* __restrict__ img, int H, int W, float x, float y,
                                       float& v, float& gx, float& gy) {
    if (interp == 0)
        sample_cubic(img, H, W, x, y, v, gx, gy);
    else
        sample_linear(img, H, W, x, y, v, gx, gy);
}

}  // namespace d2ic
