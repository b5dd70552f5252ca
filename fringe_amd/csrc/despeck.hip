// despeck.hip -- SHP-masked neighbourhood mean of amplitude, interferogram or
// coherence (src/despeck/despeck.cpp:300-433) on the GPU.
//
// Per block the reference forms two complex planes (despeck.cpp:300-352):
//   one band:          c1 = |b1|,            c2 = 1
//   two bands:         c1 = b1 conj(b2),     c2 = 1
//   two bands + corr:  c1 = b1 conj(b2),     c2 = (|b1|^2, |b2|^2)
// and, for every pixel whose centre bit is set, sums c1 and c2 over the
// masked window truncated at the block (rows) and image (columns) edges,
// row by row, in float (:391-410), then divides by Re sum c2, or by
// sqrt(Re sum c2) sqrt(Im sum c2) for coherence (:413-426).  The kernels keep
// every float operation of that sequence, in the same order and without
// fused multiply-adds (contraction is switched off in both kernels), with
// correctly rounded division and square root, so the result is
// bit-identical.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "engine_internal.hpp"

namespace fringe {

// |z| as glibc's hypotf rounds it (double hypotenuse rounded once to float).
__device__ inline float despeck_abs(float2 z) {
    const double x = (double)z.x, y = (double)z.y;
    return (float)__builtin_sqrt(x * x + y * y);
}

__global__ __launch_bounds__(256) void despeck_prep(const float2* __restrict__ b1, const float2* __restrict__ b2,
                                                    int mode, long long npx, float2* __restrict__ c1,
                                                    float2* __restrict__ c2) {
#pragma clang fp contract(off)
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npx) return;
    const float2 a = b1[p];
    if (mode == 0) {
        c1[p] = make_float2(despeck_abs(a), 0.0f);
        c2[p] = make_float2(1.0f, 0.0f);
        return;
    }
    const float2 b = b2[p];
    // a * conj(b) as the float complex product: each product rounded, then
    // the sum
    const float re = a.x * b.x + a.y * b.y;
    const float im = a.x * -b.y + a.y * b.x;
    c1[p] = make_float2(re, im);
    if (mode == 2) {
        const float m1 = despeck_abs(a), m2 = despeck_abs(b);
        c2[p] = make_float2(m1 * m1, m2 * m2);
    } else {
        c2[p] = make_float2(1.0f, 0.0f);
    }
}

__global__ __launch_bounds__(256) void despeck_sum(const float2* __restrict__ c1, const float2* __restrict__ c2,
                                                   const uint32_t* __restrict__ wts, int rows, int cols, int Nx,
                                                   int Ny, int nul, int row_begin, int row_end, int coherence,
                                                   float2* __restrict__ out) {
#pragma clang fp contract(off)
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)(row_end - row_begin) * cols) return;
    const int r = row_begin + (int)(t / cols), c = (int)(t % cols);
    const long long p = (long long)r * cols + c;
    const uint32_t* w = wts + p * nul;
    float2 res = make_float2(0.0f, 0.0f);
    if (shp_get(w, shp_flat(0, 0, Nx, Ny))) {
        const int xmin = max(c - Nx, 0), xmax = min(cols - 1, c + Nx);
        const int ymin = max(r - Ny, 0), ymax = min(rows - 1, r + Ny);
        float vr = 0.0f, vi = 0.0f, sr = 0.0f, si = 0.0f;
        for (int ii = ymin; ii <= ymax; ++ii)
            for (int jj = xmin; jj <= xmax; ++jj)
                if (shp_get(w, shp_flat(ii - r, jj - c, Nx, Ny))) {
                    const long long q = (long long)ii * cols + jj;
                    const float2 a = c1[q], b = c2[q];
                    vr = vr + a.x;
                    vi = vi + a.y;
                    sr = sr + b.x;
                    si = si + b.y;
                }
        if (sr > 0.0f) {
            if (coherence) {
                if (si > 0.0f) {
                    const float d = sqrtf(sr) * sqrtf(si);
                    res = make_float2(vr / d, vi / d);
                }
            } else {
                res = make_float2(vr / sr, vi / sr);
            }
        }
    }
    out[p] = res;
}

int despeck_launch(const float2* b1, const float2* b2, const uint32_t* wts, int rows, int cols, int Nx, int Ny,
                   int mode, int row_begin, int row_end, float2* c1, float2* c2, float2* out, hipStream_t s) {
    const long long npx = (long long)rows * cols;
    if (npx > 0)
        hipLaunchKernelGGL(despeck_prep, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, s, b1, b2, mode, npx,
                           c1, c2);
    const long long nout = (long long)(row_end - row_begin) * cols;
    if (nout > 0)
        hipLaunchKernelGGL(despeck_sum, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, s, c1, c2, wts, rows,
                           cols, Nx, Ny, shp_words(Nx, Ny), row_begin, row_end, mode == 2 ? 1 : 0, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace fringe
