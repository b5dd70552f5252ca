// calamp.hip -- amplitude calibration constant of every band
// (src/calamp/calamp.cpp:180-262): the mean of |z| over the valid pixels of
// the band, a pixel being valid when |z| is non-zero, not NaN and its mask
// byte (if any) is non-zero.  |z| is rounded to float as std::abs of a
// complex<float> rounds it, then summed in double.
//
// Each workgroup reduces a grid-strided share of the band in a fixed order
// and writes one (sum, count) partial; the host adds the partials in order,
// so the result does not depend on scheduling.  (The reference's OpenMP
// reduction does: its sum is only defined to a few ulps.)
#include <hip/hip_runtime.h>

#include "engine_internal.hpp"

namespace fringe {

constexpr int kCalampBlocks = 1024;

__device__ inline float calamp_abs(float2 z) {
    if (!isfinite(z.x) || !isfinite(z.y)) {
        if (isinf(z.x) || isinf(z.y)) return INFINITY;
        return z.x + z.y;   // NaN
    }
    const double x = (double)z.x, y = (double)z.y;
    return (float)__builtin_sqrt(x * x + y * y);
}

__global__ __launch_bounds__(256) void calamp_partial(const float2* __restrict__ z, const uint8_t* __restrict__ mask,
                                                      long long n, double2* __restrict__ partial) {
    __shared__ double ss[256], sc[256];
    double s = 0.0, c = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        double a = (double)calamp_abs(z[i]);
        a = isnan(a) ? 0.0 : a;
        const bool valid = (a != 0.0) && (!mask || mask[i] > 0);
        if (valid) {
            s += a;
            c += 1.0;
        }
    }
    ss[threadIdx.x] = s;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            ss[threadIdx.x] += ss[threadIdx.x + w];
            sc[threadIdx.x] += sc[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(ss[0], sc[0]);
}

int calamp_partials() { return kCalampBlocks; }

int calamp_launch(const float2* z, const uint8_t* mask, long long n, double2* partial, hipStream_t s) {
    hipLaunchKernelGGL(calamp_partial, dim3(kCalampBlocks), dim3(256), 0, s, z, mask, n, partial);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace fringe
