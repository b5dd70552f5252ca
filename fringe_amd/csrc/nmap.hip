// nmap.hip -- SHP (statistically homogeneous pixel) map on gfx950.
//
// Semantics follow src/nmap/nmap.cpp:204-233 and :345-473 with the
// race-free definition of the sweep (every pixel evaluates its own full
// window and writes only its own bits and count, as nmap_cuda.cu:259-334
// does): bits and counts are bit-exact against that definition.
//
// Stage 1 (nmap_prep): per pixel, amplitude |z| / alpha_b for every band,
// validity, and an ascending sort held in registers (bitonic network).
// Stage 2 (nmap_net for n <= 64, nmap_sweep above): per pixel, the
// two-sample test against every valid in-image neighbour of the forward half
// of the window (dy > 0, or dy = 0 and dx >= 0), as the reference's sweep
// visits pairs (nmap.cpp:404-457); the bits go to a workspace.  Stage 3
// (nmap_mirror) assembles every output pixel's full mask: bit (dy, dx) of
// the backward half is bit (-dy, -dx) of neighbour p + (dy, dx), i.e. the
// same test of the same ordered pair, so each pair is tested once.  KS2: the statistic as an integer count k = D*n,
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

// One thread per pixel of the whole tile.  amp_sorted is date-planar,
// [rank][pixel], so neighbouring threads read and write neighbouring words.
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
#pragma unroll
    for (int b = 0; b < NP; ++b)
        if (b < n) amp_sorted[(long long)b * npx + p] = a[b];
}

// Integer KS walk over two ascending arrays of length n (element i at
// x[i * st]), stepping exactly as KS2sample::test does
// (KS2sample.hpp:105-138): one element on strict order, every copy of a tied
// value on both sides on a tie, the running difference sampled after each
// step while both sides have elements left.  Returns k = D * n.
__device__ inline int ks_count(const float* __restrict__ a, const float* __restrict__ b, int n,
                               long long st) {
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
            while (ia < n && a[ia * st] == v) { --d; ++ia; }
            while (ib < n && b[ib * st] == v) { ++d; ++ib; }
        }
        if (ia >= n || ib >= n) break;
        x = a[ia * st];
        y = b[ib * st];
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
                                   long long st, const double* __restrict__ T) {
    int ia = 0, ib = 0, m = 0;
    double S = 0.0;
    float x = a[0], y = b[0];
    for (int j = 0; j < 2 * n - 1; ++j) {
        const bool take_a = (ia < n && ib < n) ? (x < y) : (ia < n);
        if (take_a) {
            ++ia;
            ++m;
            x = ia < n ? a[ia * st] : x;
        } else {
            ++ib;
            y = ib < n ? b[ib * st] : y;
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

__device__ inline double ad2_standardise(double S, int n, double sigma) {
    double A = 0.0;
    A = A + S / (n * 1.0);
    A = A + S / (n * 1.0);
    A = A / (2.0 * n);
    return (A - 1) / sigma;
}

__device__ inline bool ks_similar(int k, const NmapPlan& P) {
    const uint32_t tw = k < 32 ? P.table.bits[0] : k < 64 ? P.table.bits[1] : k < 96 ? P.table.bits[2]
                        : k < 128 ? P.table.bits[3] : P.table.bits[4];
    return ((tw >> (k & 31)) & 1u) != 0u;
}

// ---------------------------------------------------------------------------
// n <= 64: merge network in registers.  Positive floats order like their bit
// patterns, so key = bits << 1 | tag orders by value and, among equal values,
// by tag.  The pixel's own keys stay in registers; a neighbour's keys are read
// coalesced (date-planar layout) and the two sorted runs are merged by a
// bitonic network (own keys ascending, neighbour's descending).
//   KS2: tag 1 = neighbour.  d after a prefix = #neighbour - #own; the walk of
//        KS2sample::test samples d exactly at the ends of runs of equal
//        values (inside a run of one sample d is monotone, so the samples
//        the walk takes there add nothing), so k = max |d| over positions
//        whose successor has a different value.
//   AD2: tag 1 = the test's first sample, so ties put the second sample
//        first as AD2unique::test merges them; m_j = number of tag-1 keys in
//        the first j+1 positions.
// Pads (beyond n) get keys above every finite or infinite amplitude.
// ---------------------------------------------------------------------------
// Forward-half words of the pixels of rows [fwd_row0, row_end): the
// output rows and the Ny rows above them whose forward tests reach into
// them.  Pixel (r, c) keeps words centre/32 .. nul-1 of its mask at
// F + ((r - fwd_row0) * cols + c) * nf, and its forward count (centre
// included) at fcnt[(r - fwd_row0) * cols + c].
__device__ inline int fwd_row0(const NmapPlan& P) { return max(0, P.row_begin - P.Ny); }
__device__ inline int centre_flat(const NmapPlan& P) { return P.Ny * (2 * P.Nx + 1) + P.Nx; }

template <int NP, bool AD2>
__global__ __launch_bounds__(256) void nmap_net(const NmapPlan P, uint32_t* __restrict__ F,
                                                int32_t* __restrict__ fcnt) {
    const uint32_t* __restrict__ keys = reinterpret_cast<const uint32_t*>(P.d_amp);
    const uint8_t* __restrict__ valid = P.d_valid;
    const int n = P.n, rows = P.rows, cols = P.cols, Nx = P.Nx, Ny = P.Ny, nul = P.nul;
    const long long npx = (long long)rows * cols;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int h0 = fwd_row0(P);
    if (t >= (long long)(P.row_end - h0) * cols) return;
    const int r = h0 + (int)(t / cols);
    const int c = (int)(t % cols);
    const long long p = (long long)r * cols + c;
    const int centre = centre_flat(P), fw = centre >> 5, nf = nul - fw;
    uint32_t* out = F + t * nf;   // word w of the mask at out[w - fw]
    const bool pv = valid[p] != 0;
    uint32_t A[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) A[i] = (pv && i < n) ? (keys[i * npx + p] << 1) : 0xFFFFFFFEu;
    int cnt = 0, word = fw, flat = centre;
    uint32_t cur = 0u;
    for (int dy = 0; dy <= Ny; ++dy) {
        const int rr = r + dy;
        for (int dx = (dy == 0 ? 0 : -Nx); dx <= Nx; ++dx, ++flat) {
            if ((flat >> 5) != word) {
                out[word - fw] = cur;
                cur = 0u;
                word = flat >> 5;
            }
            const int cc = c + dx;
            if (!pv || rr >= rows || cc < 0 || cc >= cols) continue;
            const long long q = (long long)rr * cols + cc;
            if (!valid[q]) continue;
            bool similar = true;
            if (q != p) {
                // own tag: KS2 0; AD2 1 (the pixel is the test's first sample)
                const uint32_t tp = AD2 ? 1u : 0u, tq = tp ^ 1u;
                uint32_t S[2 * NP];
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    S[i] = A[i] | tp;
                    S[2 * NP - 1 - i] = (i < n ? (keys[i * npx + q] << 1) : 0xFFFFFFFEu) | tq;
                }
#pragma unroll
                for (int s = NP; s >= 1; s >>= 1)
#pragma unroll
                    for (int i = 0; i < 2 * NP; ++i)
                        if ((i & s) == 0) {
                            const uint32_t x = S[i], y = S[i + s];
                            S[i] = min(x, y);
                            S[i + s] = max(x, y);
                        }
                if constexpr (AD2) {
                    // first sample = tag 1
                    int m = 0;
                    double Ssum = 0.0;
#pragma unroll
                    for (int j = 0; j < 2 * NP - 1; ++j) {
                        m += (int)(S[j] & 1u);
                        if (j < 2 * n - 1) Ssum = Ssum + P.d_ad2[j * (n + 1) + m];
                    }
                    similar = ad2_standardise(Ssum, n, P.ad2_sigma) <= P.ad2_max;
                } else {
                    // d_j = 2 (#tag-1 in the first j+1) - (j+1), taken where
                    // the amplitude changes (the end of a run of ties).
                    // Padding keys sort above every amplitude and differ from
                    // one another only in the tag, so no end falls among them,
                    // and at the last amplitude d = 0: j needs no bound.
                    int nb = 0, hi = 0, lo = 0;
#pragma unroll
                    for (int j = 0; j < 2 * NP - 1; ++j) {
                        nb += (int)(S[j] & 1u);
                        const bool end = (S[j] ^ S[j + 1]) > 1u;
                        const int x = end ? 2 * nb - (j + 1) : 0;
                        hi = max(hi, x);
                        lo = min(lo, x);
                    }
                    similar = ks_similar(max(hi, -lo), P);
                }
            }
            if (similar) {
                cur |= (1u << (flat & 31));
                ++cnt;
            }
        }
    }
    out[word - fw] = cur;
    for (int w = word + 1; w < nul; ++w) out[w - fw] = 0u;
    fcnt[t] = cnt;
}

// n > 64: one thread per output pixel walks the sorted arrays in memory.
// The reference sweeps half the window and tests (p, q) with p first for
// q = p + (dy, dx), dy > 0 or dy = 0 < dx (nmap.cpp:402-457); the other half
// of a pixel's window is therefore tested with the neighbour first.
template <bool AD2>
__global__ __launch_bounds__(256) void nmap_sweep(const NmapPlan P, uint32_t* __restrict__ F,
                                                  int32_t* __restrict__ fcnt) {
    const float* __restrict__ amp_sorted = P.d_amp;
    const uint8_t* __restrict__ valid = P.d_valid;
    const int n = P.n, rows = P.rows, cols = P.cols, Nx = P.Nx, Ny = P.Ny, nul = P.nul;
    const long long npx = (long long)rows * cols;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int h0 = fwd_row0(P);
    if (t >= (long long)(P.row_end - h0) * cols) return;
    const int r = h0 + (int)(t / cols);
    const int c = (int)(t % cols);
    const long long p = (long long)r * cols + c;
    const int centre = centre_flat(P), fw = centre >> 5, nf = nul - fw;
    uint32_t* out = F + t * nf;   // word w of the mask at out[w - fw]
    const bool pv = valid[p] != 0;
    const float* a = amp_sorted + p;
    // Window positions are visited in bit order (flat = centre, centre + 1,
    // ...), so the word being filled only ever advances by one; each
    // finished word is stored once.
    int cnt = 0, word = fw, flat = centre;
    uint32_t cur = 0u;
    for (int dy = 0; dy <= Ny; ++dy) {
        const int rr = r + dy;
        for (int dx = (dy == 0 ? 0 : -Nx); dx <= Nx; ++dx, ++flat) {
            if ((flat >> 5) != word) {
                out[word - fw] = cur;
                cur = 0u;
                word = flat >> 5;
            }
            const int cc = c + dx;
            if (!pv || rr >= rows || cc < 0 || cc >= cols) continue;
            const long long q = (long long)rr * cols + cc;
            if (!valid[q]) continue;
            bool similar = true;
            if (q != p) {
                const float* b = amp_sorted + q;
                if constexpr (AD2) {   // the pixel is the test's first sample
                    const double S = ad2_inner(a, b, n, npx, P.d_ad2);
                    similar = ad2_standardise(S, n, P.ad2_sigma) <= P.ad2_max;
                } else {
                    similar = ks_similar(ks_count(a, b, n, npx), P);
                }
            }
            if (similar) {
                cur |= (1u << (flat & 31));
                ++cnt;
            }
        }
    }
    out[word - fw] = cur;
    for (int w = word + 1; w < nul; ++w) out[w - fw] = 0u;
    fcnt[t] = cnt;
}

// Full masks of the output rows.  Backward position flat (< centre) of
// pixel p holds the forward bit W-1-flat of q = p + (dy, dx), the pair's
// one test; q lies in rows [fwd_row0, row_end) whenever it is in the image.
__global__ __launch_bounds__(256) void nmap_mirror(const NmapPlan P, const uint32_t* __restrict__ F,
                                                   const int32_t* __restrict__ fcnt, uint32_t* __restrict__ wts,
                                                   int32_t* __restrict__ count) {
    const int cols = P.cols, Nx = P.Nx, Ny = P.Ny, nul = P.nul;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)(P.row_end - P.row_begin) * cols) return;
    const int r = P.row_begin + (int)(t / cols);
    const int c = (int)(t % cols);
    const int h0 = fwd_row0(P);
    const int centre = centre_flat(P), fw = centre >> 5, nf = nul - fw;
    const int last = (2 * Ny + 1) * (2 * Nx + 1) - 1;
    const long long fp = (long long)(r - h0) * cols + c;
    uint32_t* out = wts + ((long long)r * cols + c) * nul;
    int cnt = fcnt[fp];
    int word = 0, flat = 0;
    uint32_t cur = 0u;
    for (int dy = -Ny; dy <= 0; ++dy) {
        const int rr = r + dy;
        for (int dx = -Nx; dx <= (dy == 0 ? -1 : Nx); ++dx, ++flat) {
            if ((flat >> 5) != word) {
                out[word] = cur;
                cur = 0u;
                word = flat >> 5;
            }
            const int cc = c + dx;
            if (rr < 0 || cc < 0 || cc >= cols) continue;
            const int mf = last - flat;
            const uint32_t qw = F[((long long)(rr - h0) * cols + cc) * nf + (mf >> 5) - fw];
            if ((qw >> (mf & 31)) & 1u) {
                cur |= (1u << (flat & 31));
                ++cnt;
            }
        }
    }
    // flat == centre: the rest comes from the pixel's own forward words
    if (fw != word) {
        out[word] = cur;
        cur = 0u;
    }
    const uint32_t* mine = F + fp * nf;
    out[fw] = cur | mine[0];
    for (int i = 1; i < nf; ++i) out[fw + i] = mine[i];
    count[(long long)r * cols + c] = cnt;
}

int nmap_fwd_words(int Nx, int Ny) { return shp_words(Nx, Ny) - ((Ny * (2 * Nx + 1) + Nx) >> 5); }

template <int NP>
static void launch_prep(const NmapPlan& P, const float2* slc, const uint8_t* mask, hipStream_t s) {
    const long long npx = (long long)P.rows * P.cols;
    const int threads = 256;
    const long long blocks = (npx + threads - 1) / threads;
    hipLaunchKernelGGL(nmap_prep<NP>, dim3((unsigned)blocks), dim3(threads), 0, s, P, slc, mask);
}

template <int NP>
static void launch_net(bool ad2, dim3 g, dim3 blk, const NmapPlan& P, hipStream_t s) {
    if (ad2) hipLaunchKernelGGL((nmap_net<NP, true>), g, blk, 0, s, P, P.d_fwd, P.d_fcnt);
    else hipLaunchKernelGGL((nmap_net<NP, false>), g, blk, 0, s, P, P.d_fwd, P.d_fcnt);
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
        // forward halves of the output rows and of the Ny rows above them
        const int h0 = P.row_begin - P.Ny > 0 ? P.row_begin - P.Ny : 0;
        const long long nfwd = (long long)(P.row_end - h0) * P.cols;
        StageScope sc(STAGE_NMAP_KS, s);
        if (P.method == 1) {
            const int nt = (2 * P.n - 1) * (P.n + 1);
            hipLaunchKernelGGL(ad2_terms, dim3((nt + 255) / 256), dim3(256), 0, s, P.d_ad2, P.n);
        }
        const bool ad2 = P.method == 1;
        const dim3 g((unsigned)((nfwd + threads - 1) / threads)), blk(threads);
        if (P.n <= 8) launch_net<8>(ad2, g, blk, P, s);
        else if (P.n <= 16) launch_net<16>(ad2, g, blk, P, s);
        else if (P.n <= 32) launch_net<32>(ad2, g, blk, P, s);
        else if (P.n <= 64) launch_net<64>(ad2, g, blk, P, s);
        else if (ad2) hipLaunchKernelGGL(nmap_sweep<true>, g, blk, 0, s, P, P.d_fwd, P.d_fcnt);
        else hipLaunchKernelGGL(nmap_sweep<false>, g, blk, 0, s, P, P.d_fwd, P.d_fcnt);
        hipLaunchKernelGGL(nmap_mirror, dim3((unsigned)((nout + threads - 1) / threads)), dim3(threads), 0, s,
                           P, (const uint32_t*)P.d_fwd, (const int32_t*)P.d_fcnt, wts, count);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace fringe
