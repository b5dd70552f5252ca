// adjust.hip -- datum adjustment of sequential mini-stacks on the GPU.
//
// adjustMiniStacks.py:157-228 expresses the adjustment as GDAL VRT pixel
// functions: "mul" for wrapped phases (mini-stack phase x datum phase of the
// mini-stack, CFloat32) and "sum" for unwrapped phases (Float32).  GDAL
// evaluates both in double and rounds the result to the band type; the
// kernels do the same, so the materialised rasters equal what reading the
// reference's VRT would return.  Elementwise and HBM-bound: 24 B (wrapped) or
// 12 B (unwrapped) per pixel.
#include <hip/hip_runtime.h>

#include "engine_internal.hpp"
#include "fringe_amd.h"

namespace fringe {

__global__ __launch_bounds__(256) void adjust_wrapped_kernel(const float2* __restrict__ a,
                                                             const float2* __restrict__ b,
                                                             float2* __restrict__ out, long long npx) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < npx;
         i += (long long)gridDim.x * blockDim.x) {
        const float2 x = a[i], y = b[i];
        // products of floats are exact in double; one rounding per sum, one
        // to float
        const double re = (double)x.x * (double)y.x - (double)x.y * (double)y.y;
        const double im = (double)x.x * (double)y.y + (double)x.y * (double)y.x;
        out[i] = make_float2((float)re, (float)im);
    }
}

__global__ __launch_bounds__(256) void adjust_unwrapped_kernel(const float* __restrict__ a,
                                                               const float* __restrict__ b,
                                                               float* __restrict__ out, long long npx) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < npx;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = (float)((double)a[i] + (double)b[i]);
}

static unsigned adjust_blocks(long long npx) {
    const long long b = (npx + 255) / 256;
    return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

}  // namespace fringe

using namespace fringe;

extern "C" {

int fringe_adjust_wrapped(const void* d_mini, const void* d_datum, void* d_out, long long npx, void* stream) {
    if (!d_mini || !d_datum || !d_out || npx < 0) return FRINGE_ERR_ARGUMENT;
    if (npx == 0) return 0;
    hipLaunchKernelGGL(adjust_wrapped_kernel, dim3(adjust_blocks(npx)), dim3(256), 0, (hipStream_t)stream,
                       (const float2*)d_mini, (const float2*)d_datum, (float2*)d_out, npx);
    if (hipGetLastError() != hipSuccess) {
        set_last_error("fringe_adjust_wrapped: launch failed");
        return FRINGE_ERR_DEVICE;
    }
    return 0;
}

int fringe_adjust_unwrapped(const float* d_mini, const float* d_datum, float* d_out, long long npx,
                            void* stream) {
    if (!d_mini || !d_datum || !d_out || npx < 0) return FRINGE_ERR_ARGUMENT;
    if (npx == 0) return 0;
    hipLaunchKernelGGL(adjust_unwrapped_kernel, dim3(adjust_blocks(npx)), dim3(256), 0, (hipStream_t)stream,
                       d_mini, d_datum, d_out, npx);
    if (hipGetLastError() != hipSuccess) {
        set_last_error("fringe_adjust_unwrapped: launch failed");
        return FRINGE_ERR_DEVICE;
    }
    return 0;
}

}  // extern "C"
