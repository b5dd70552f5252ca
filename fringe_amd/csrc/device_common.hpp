// device_common.hpp -- complex helpers, SHP bit layout and lane-group
// primitives shared by the nmap and phase-linking kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fringe {

// ---- SHP bitmask layout (ulongmask.hpp:55-94) -----------------------------
// Window offset (dy, dx), dy in [-Ny, Ny], dx in [-Nx, Nx] maps to
// flat = (dy+Ny)(2Nx+1) + (dx+Nx); word flat>>5, bit flat&31 (LSB first).
__host__ __device__ inline int shp_flat(int dy, int dx, int Nx, int Ny) {
    return (dy + Ny) * (2 * Nx + 1) + (dx + Nx);
}
__host__ __device__ inline int shp_words(int Nx, int Ny) {
    return ((2 * Ny + 1) * (2 * Nx + 1) + 31) / 32;
}
__host__ __device__ inline bool shp_get(const uint32_t* w, int flat) {
    return (w[flat >> 5] >> (flat & 31)) & 1u;
}

// ---- complex arithmetic on (T, T) pairs ----------------------------------
template <typename T> struct cx {
    T re, im;
};
template <typename T> __device__ inline cx<T> cmake(T r, T i) { return cx<T>{r, i}; }
template <typename T> __device__ inline cx<T> cadd(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T> __device__ inline cx<T> csub(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T> __device__ inline cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// a * conj(b)
template <typename T> __device__ inline cx<T> cmulc(cx<T> a, cx<T> b) {
    return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
template <typename T> __device__ inline cx<T> cconj(cx<T> a) { return {a.re, -a.im}; }
template <typename T> __device__ inline cx<T> cscale(cx<T> a, T s) { return {a.re * s, a.im * s}; }
template <typename T> __device__ inline T cabs2(cx<T> a) { return a.re * a.re + a.im * a.im; }
// acc += a * conj(b)
template <typename T> __device__ inline void cfma_c(cx<T>& acc, cx<T> a, cx<T> b) {
    acc.re = fma(a.re, b.re, fma(a.im, b.im, acc.re));
    acc.im = fma(a.im, b.re, fma(-a.re, b.im, acc.im));
}
// acc -= a * conj(b), as four FMAs
template <typename T> __device__ inline void cfms_c(cx<T>& acc, cx<T> a, cx<T> b) {
    acc.re = fma(-a.re, b.re, fma(-a.im, b.im, acc.re));
    acc.im = fma(-a.im, b.re, fma(a.re, b.im, acc.im));
}
// acc += a * b
template <typename T> __device__ inline void cfma(cx<T>& acc, cx<T> a, cx<T> b) {
    acc.re = fma(a.re, b.re, fma(-a.im, b.im, acc.re));
    acc.im = fma(a.re, b.im, fma(a.im, b.re, acc.im));
}
template <typename T> __device__ inline cx<T> crecip(cx<T> a) {
    T s = T(1) / (a.re * a.re + a.im * a.im);
    return {a.re * s, -a.im * s};
}

// ---- lane groups ---------------------------------------------------------
// G <= 64: a wave of 64 lanes is split into 64/G groups of G lanes, each
// working on one pixel, and group reductions use cross-lane shuffles limited
// to the group's lanes.  G > 64: a workgroup of G/64 waves works on one
// pixel; each wave reduces by shuffles and the G/64 partials meet in LDS
// (every lane of the workgroup must take part: the barriers are block-wide).
// One 32-byte LDS scratch serves every multi-wave reduction of a kernel
// (partials travel as double, exact for float and double operands), so the
// reductions add 32 bytes of static LDS to a kernel whatever their number.
__device__ inline double* greduce_scratch() {
    __shared__ double part[4];
    return part;
}
template <int G, typename T, typename F> __device__ inline T greduce(T x, F f) {
    static_assert(G <= 256, "at most four waves per group");
    const int W = G > 64 ? 64 : G;
#pragma unroll
    for (int o = W / 2; o >= 1; o >>= 1) x = f(x, __shfl_xor(x, o, W));
    if constexpr (G > 64) {
        static_assert(sizeof(T) <= sizeof(double), "floating-point partials");
        double* part = greduce_scratch();
        __syncthreads();   // the previous reduction's reads are done
        if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = (double)x;
        __syncthreads();
        x = (T)part[0];
#pragma unroll
        for (int w = 1; w < G / 64; ++w) x = f(x, (T)part[w]);
    }
    return x;
}
template <int G, typename T> __device__ inline T gsum(T x) {
    return greduce<G>(x, [](T a, T b) { return a + b; });
}
template <int G, typename T> __device__ inline T gmax(T x) {
    return greduce<G>(x, [](T a, T b) { return max(a, b); });
}
template <int G, typename T> __device__ inline T gmin(T x) {
    return greduce<G>(x, [](T a, T b) { return min(a, b); });
}
template <int G, typename T> __device__ inline cx<T> gsum(cx<T> x) {
    return cx<T>{gsum<G>(x.re), gsum<G>(x.im)};
}
// Bits of a wave ballot that belong to the calling lane's group.
template <int G> __device__ inline uint64_t group_ballot(bool p) {
    uint64_t b = __ballot(p);
    if (G == 64) return b;
    const int grp = (threadIdx.x & 63) / G;
    const uint64_t m = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
    return (b >> (grp * G)) & m;
}

// Host-visible error record for kernels that detect an impossible launch
// shape; every kernel checks its arguments and returns early instead of
// touching memory out of bounds.
}  // namespace fringe
