// nmap.hip -- SHP (statistically homogeneous pixel) map on gfx950.
//
// Semantics follow src/nmap/nmap.cpp:204-233 and :345-473 with the
// race-free definition of the sweep (every pixel evaluates its own full
// window and writes only its own bits and count, as nmap_cuda.cu:259-334
// does): bits and counts are bit-exact against that definition.
//
// Stage 1 (nmap_prep): per pixel, amplitude |z| / alpha_b for every band,
// validity, and an ascending sort held in registers (bitonic network).
// Stage 2 (nmap_sweep): per pixel, the two-sample test against every valid
// in-image neighbour.  KS2: the statistic as an integer count k = D*n,
// decided through a host-built table (k -> similar) instead of evaluating the
// Kolmogorov series on the device.  AD2: the standardised Anderson-Darling
// statistic in double, bit-identical to the host's, compared with the largest
// statistic whose interpolated p-value still reaches the threshold (found on
// the host, so no exp/log runs on the device).
#include "device_common.hpp"
#include "engine_internal.hpp"
#include "timing.hpp"

namespace fringe {

// |z| exactly as glibc's hypotf computes it for finite input (the value
// std::abs(std::complex<float>) returns on the host): the double-precision
// hypotenuse rounded once to float.  Non-finite parts give +inf when either
// part is infinite, NaN otherwise.
__device__ inline float amp_hypot(float x, float y) {
    if (!isfinite(x) || !isfinite(y)) {
        if (isinf(x) || isinf(y)) return INFINITY;
        return x + y;
    }
    double dx = (double)x, dy = (double)y;
    return (float)__builtin_sqrt(dx * dx + dy * dy);
}

template <int NP> __device__ inline void bitonic_sort(float (&a)[NP]) {
#pragma unroll
    for (int k = 2; k <= NP; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    const bool up = ((i & k) == 0);
                    const float x = a[i], y = a[l];
                    const float lo = fminf(x, y), hi = fmaxf(x, y);
                    a[i] = up ? lo : hi;
                    a[l] = up ? hi : lo;
                }
            }
        }
    }
}

// One thread per pixel of the whole tile.  amp_sorted is [pixel][n].
template <int NP>
__global__ __launch_bounds__(256) void nmap_prep(const NmapPlan P, const float2* __restrict__ slc,
                                                 const uint8_t* __restrict__ mask) {
    const long long npx = (long long)P.rows * P.cols;
    const int n = P.n;
    float* __restrict__ amp_sorted = P.d_amp;
    uint8_t* __restrict__ valid = P.d_valid;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npx) return;
    float a[NP];
    bool ok = mask ? (mask[p] != 0) : true;
#pragma unroll
    for (int b = 0; b < NP; ++b) {
        a[b] = INFINITY;
        if (b < n) {
            const float2 z = slc[(long long)b * npx + p];
            const float val = (float)((double)amp_hypot(z.x, z.y) / P.alpha[b]);
            ok = ok && (val != 0.0f) && !isnan(val);
            a[b] = val;
        }
    }
    valid[p] = ok ? 1 : 0;
    if (!ok) return;
    bitonic_sort<NP>(a);
    float* out = amp_sorted + p * n;
#pragma unroll
    for (int b = 0; b < NP; ++b)
        if (b < n) out[b] = a[b];
}

// Integer KS walk over two ascending arrays of length n, stepping exactly as
// KS2sample::test does (KS2sample.hpp:105-138): one element on strict
// order, every copy of a tied value on both sides on a tie, the running
// difference sampled after each step while both sides have elements left.
// Returns k = D * n.
__device__ inline int ks_count(const float* __restrict__ a, const float* __restrict__ b, int n) {
    int ia = 0, ib = 0, d = 0, dmax = 0;
    float x = a[0], y = b[0];
    for (int it = 0; it < 2 * n; ++it) {
        if (x < y) {
            --d;
            ++ia;
        } else if (x > y) {
            ++d;
            ++ib;
        } else {
            const float v = x;
            while (ia < n && a[ia] == v) { --d; ++ia; }
            while (ib < n && b[ib] == v) { ++d; ++ib; }
        }
        if (ia >= n || ib >= n) break;
        x = a[ia];
        y = b[ib];
        dmax = max(dmax, abs(d));
    }
    return max(dmax, abs(d));
}

// Anderson-Darling inner sum of AD2unique::test (AD2unique.hpp:206-351) for
// ascending a (first sample) and b: merge with ties taken from b first and
// add, for every merged position j < 2n-1, the term of (j, m_j) where m_j
// counts the a values among the first j+1.  The terms are tabulated once per
// call (ad2_terms), so the double sum adds the same values in the same order.
__device__ inline double ad2_inner(const float* __restrict__ a, const float* __restrict__ b, int n,
                                   const double* __restrict__ T) {
    int ia = 0, ib = 0, m = 0;
    double S = 0.0;
    float x = a[0], y = b[0];
    for (int j = 0; j < 2 * n - 1; ++j) {
        const bool take_a = (ia < n && ib < n) ? (x < y) : (ia < n);
        if (take_a) {
            ++ia;
            ++m;
            x = ia < n ? a[ia] : x;
        } else {
            ++ib;
            y = ib < n ? b[ib] : y;
        }
        S = S + T[j * (n + 1) + m];
    }
    return S;
}

// term(j, m) = (2n m - n (j+1))^2 / ((j+1)(2n - j - 1)); every intermediate
// is an exact integer in double, so the one rounding is the division's.
__global__ void ad2_terms(double* __restrict__ T, int n) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (2 * n - 1) * (n + 1)) return;
    const int j = idx / (n + 1), m = idx % (n + 1);
    const double nsum = 2.0 * n, bj = j + 1;
    const double t = nsum * m - (double)n * bj;
    T[idx] = t * t / (bj * (nsum - bj));
}

// One thread per output pixel (rows [row_begin, row_end) of the tile).
// The reference sweeps half the window and tests (p, q) with p first for
// q = p + (dy, dx), dy > 0 or dy = 0 < dx (nmap.cpp:402-457); the other half
// of a pixel's window is therefore tested with the neighbour first.
template <bool AD2>
__global__ __launch_bounds__(256) void nmap_sweep(const NmapPlan P, uint32_t* __restrict__ wts,
                                                  int32_t* __restrict__ count) {
    const float* __restrict__ amp_sorted = P.d_amp;
    const uint8_t* __restrict__ valid = P.d_valid;
    const int n = P.n, rows = P.rows, cols = P.cols, Nx = P.Nx, Ny = P.Ny, nul = P.nul;
    const int row_begin = P.row_begin, row_end = P.row_end;
    // decision bits k -> similar, held in scalars and picked by a select chain
    const uint32_t t0 = P.table.bits[0], t1 = P.table.bits[1], t2 = P.table.bits[2],
                   t3 = P.table.bits[3], t4 = P.table.bits[4];
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nout = (long long)(row_end - row_begin) * cols;
    if (t >= nout) return;
    const int r = row_begin + (int)(t / cols);
    const int c = (int)(t % cols);
    const long long p = (long long)r * cols + c;
    uint32_t* out = wts + p * nul;
    const bool pv = valid[p] != 0;
    const float* a = amp_sorted + p * n;
    // Window positions are visited in bit order (flat = 0, 1, 2, ...), so
    // the word being filled only ever advances by one; each finished word is
    // stored once.
    int cnt = 0, word = 0, flat = 0;
    uint32_t cur = 0u;
    for (int dy = -Ny; dy <= Ny; ++dy) {
        const int rr = r + dy;
        for (int dx = -Nx; dx <= Nx; ++dx, ++flat) {
            if ((flat >> 5) != word) {
                out[word] = cur;
                cur = 0u;
                word = flat >> 5;
            }
            const int cc = c + dx;
            if (!pv || rr < 0 || rr >= rows || cc < 0 || cc >= cols) continue;
            const long long q = (long long)rr * cols + cc;
            if (!valid[q]) continue;
            bool similar = true;
            if (q != p) {
                const float* b = amp_sorted + q * n;
                if constexpr (AD2) {
                    const bool fwd = dy > 0 || (dy == 0 && dx > 0);
                    const double S = ad2_inner(fwd ? a : b, fwd ? b : a, n, P.d_ad2);
                    double A = 0.0;
                    A = A + S / (n * 1.0);
                    A = A + S / (n * 1.0);
                    A = A / (2.0 * n);
                    similar = (A - 1) / P.ad2_sigma <= P.ad2_max;
                } else {
                    const int k = ks_count(a, b, n);
                    const uint32_t tw = k < 32 ? t0 : k < 64 ? t1 : k < 96 ? t2 : k < 128 ? t3 : t4;
                    similar = ((tw >> (k & 31)) & 1u) != 0u;
                }
            }
            if (similar) {
                cur |= (1u << (flat & 31));
                ++cnt;
            }
        }
    }
    out[word] = cur;
    for (int w = word + 1; w < nul; ++w) out[w] = 0u;
    count[p] = cnt;
}

template <int NP>
static void launch_prep(const NmapPlan& P, const float2* slc, const uint8_t* mask, hipStream_t s) {
    const long long npx = (long long)P.rows * P.cols;
    const int threads = 256;
    const long long blocks = (npx + threads - 1) / threads;
    hipLaunchKernelGGL(nmap_prep<NP>, dim3((unsigned)blocks), dim3(threads), 0, s, P, slc, mask);
}

int nmap_launch(const NmapPlan& P, const float2* slc, const uint8_t* mask, uint32_t* wts,
                int32_t* count, hipStream_t s) {
    {
        StageScope sc(STAGE_NMAP_PREP, s);
        if (P.n <= 8) launch_prep<8>(P, slc, mask, s);
        else if (P.n <= 16) launch_prep<16>(P, slc, mask, s);
        else if (P.n <= 32) launch_prep<32>(P, slc, mask, s);
        else if (P.n <= 64) launch_prep<64>(P, slc, mask, s);
        else launch_prep<128>(P, slc, mask, s);
    }
    const long long nout = (long long)(P.row_end - P.row_begin) * P.cols;
    if (nout > 0) {
        const int threads = 256;
        const long long blocks = (nout + threads - 1) / threads;
        StageScope sc(STAGE_NMAP_KS, s);
        if (P.method == 1) {
            const int nt = (2 * P.n - 1) * (P.n + 1);
            hipLaunchKernelGGL(ad2_terms, dim3((nt + 255) / 256), dim3(256), 0, s, P.d_ad2, P.n);
            hipLaunchKernelGGL(nmap_sweep<true>, dim3((unsigned)blocks), dim3(threads), 0, s, P, wts, count);
        } else {
            hipLaunchKernelGGL(nmap_sweep<false>, dim3((unsigned)blocks), dim3(threads), 0, s, P, wts, count);
        }
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace fringe
