// gram_mfma.hip -- masked-neighbourhood coherence on the MFMA units (gfx950).
//
// For a pixel with compacted SHP neighbour list k = 1..K (positions whose
// mask bit is set and which lie inside the tile, evd.cpp:530-564), the
// complex Gram G = sum_k z_k z_k^H is computed as the real Gram P = Y^T Y of
// Y = [Re z | Im z] (K x 2n):
//     Re G_ij = P(re_i, re_j) + P(im_i, im_j)
//     Im G_ij = P(im_i, re_j) - P(re_i, im_j)
// Each sample is split once into a 16-bit hi + lo pair (hi = h(x), lo =
// h(x - hi)) and P is accumulated in FP32 by v_mfma_f32_32x32x16_{f16,bf16}
// over hi*hi + hi*lo + lo*hi.
//   * fp16 (default): every date is first scaled by an exact power of two so
//     that its largest |Re|, |Im| over the tile lies in [2^14, 2^15); the pair
//     then carries ~22 significant bits.  The coherence
//     Gamma_ij = G_ij / sqrt(G_ii G_jj) is invariant to per-date scaling, so
//     the scale never has to be undone.
//   * bf16: the pair carries ~16 bits and keeps the FP32 exponent range, so
//     no scaling is applied.
//
// Data flow:
//   pack_hilo : the band-planar stack slc[t][row][col] is rewritten once as
//               position-major packed words pk[pixel][re_0, im_0, re_1,
//               im_1, ...] with (hi | lo << 16); one extra all-zero record at
//               index rows*cols pads neighbour lists.
//   gram_mfma : one wave per pixel at a time.  The wave compacts the mask
//               into a list of pixel indices (padded to a multiple of 16) in
//               LDS; per 16 positions lane l gathers component (l % 32) of 8
//               positions for every 32-wide component block -- lanes 0..31
//               read one contiguous 128-byte record slice, lanes 32..63
//               another -- and issues 3 MFMAs per block pair.  The upper
//               block triangle of P is folded into Gamma (the transposed
//               cross block goes through a per-wave LDS scratch) and written.
//               Above 32 dates (gram_mfma_pairs) the wave walks the pairs of
//               32-date blocks one at a time instead of holding every block
//               (at 33-64 dates 9-14 % faster than gram_mfma holding both).
#include <algorithm>

#include "device_common.hpp"
#include "engine_internal.hpp"
#include "fringe_amd.h"

namespace fringe {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t u32x4;

// Round-to-nearest-even float -> bf16 bits (finite input).
__device__ inline uint32_t bf16_rne(float x) {
    const uint32_t u = __float_as_uint(x);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__device__ inline uint32_t split_hi_lo_bf16(float x) {
    const uint32_t hi = bf16_rne(x);
    const float rest = x - __uint_as_float(hi << 16);
    const uint32_t lo = bf16_rne(rest);
    return hi | (lo << 16);
}
__device__ inline uint32_t split_hi_lo_f16(float x) {
    const _Float16 hi = (_Float16)x;
    const _Float16 lo = (_Float16)(x - (float)hi);
    return (uint32_t)__builtin_bit_cast(uint16_t, hi) | ((uint32_t)__builtin_bit_cast(uint16_t, lo) << 16);
}

// ---- per-date largest |Re|, |Im| over the tile ----------------------------
// Non-negative floats order like their bit patterns, so the maximum is kept
// as an unsigned word with atomicMax.
__global__ void date_absmax_clear(uint32_t* dmax, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dmax[t] = 0u;
}
__global__ __launch_bounds__(256) void date_absmax(const float2* __restrict__ slc, long long npx,
                                                   uint32_t* __restrict__ dmax) {
    const int t = blockIdx.y;
    const float2* __restrict__ band = slc + (long long)t * npx;
    const long long stride = (long long)gridDim.x * blockDim.x;
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float m = 0.0f;
    // Largest finite |Re|, |Im|: an infinite or NaN sample makes only the
    // coherence of the windows that hold it non-finite (as in FP64), and must
    // not take the scale of the whole date with it.  Four independent loads
    // in flight per thread; the maximum does not depend on the order.
    auto fin_abs = [](float2 z) {
        const float a = fmaxf(fabsf(z.x), fabsf(z.y));   // NaN parts dropped
        return (a < INFINITY && z.x == z.x && z.y == z.y) ? a : 0.0f;
    };
    for (; p + 3 * stride < npx; p += 4 * stride) {
        float2 z[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) z[k] = band[p + k * stride];
#pragma unroll
        for (int k = 0; k < 4; ++k) m = fmaxf(m, fin_abs(z[k]));
    }
    for (; p < npx; p += stride) m = fmaxf(m, fin_abs(band[p]));
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    // one atomic per workgroup: every date's word shares a cache line, and
    // atomics on one line serialise
    __shared__ float wm[4];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(dmax + t, __float_as_uint(fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]))));
}
// Power-of-two exponent k with max * 2^k in [2^14, 2^15); 0 when the date is
// all zero or not finite.
__device__ inline int date_scale_exp(uint32_t bits) {
    const float m = __uint_as_float(bits);
    if (!(m > 0.0f) || !isfinite(m)) return 0;
    int e;
    frexpf(m, &e);          // m in [2^(e-1), 2^e)
    return min(15 - e, 126);
}

// ---- pack: [t][pixel] complex64 -> [pixel][2n] (hi|lo) words -------------
// 64 pixels per workgroup; the band-planar rows are read coalesced and the
// records written contiguously through an LDS transpose.
template <bool F16>
__global__ __launch_bounds__(256) void pack_hilo(const float2* __restrict__ slc, int n,
                                                 long long npx, const uint32_t* __restrict__ dmax,
                                                 uint32_t* __restrict__ pk) {
    extern __shared__ uint32_t st[];   // [64][2n]
    const long long p0 = (long long)blockIdx.x * 64;
    const int CS = 2 * n;
    for (int e = threadIdx.x; e < 64 * n; e += blockDim.x) {
        const int t = e / 64, i = e % 64;
        const long long p = p0 + i;
        float2 z = make_float2(0.0f, 0.0f);
        if (p < npx) z = slc[(long long)t * npx + p];
        if (F16) {
            const int k = date_scale_exp(dmax[t]);
            st[i * CS + 2 * t] = split_hi_lo_f16(ldexpf(z.x, k));
            st[i * CS + 2 * t + 1] = split_hi_lo_f16(ldexpf(z.y, k));
        } else {
            st[i * CS + 2 * t] = split_hi_lo_bf16(z.x);
            st[i * CS + 2 * t + 1] = split_hi_lo_bf16(z.y);
        }
    }
    __syncthreads();
    const long long words = (long long)min((long long)64, npx - p0) * CS;
    for (long long e = threadIdx.x; e < words; e += blockDim.x) pk[p0 * CS + e] = st[e];
    if (blockIdx.x == 0)
        for (int e = threadIdx.x; e < CS; e += blockDim.x) pk[npx * CS + e] = 0u;
}

// The staged kernel reads a pixel's mask words from lanes 0..15, two words
// per 64-position pass: windows of up to 512 positions.
constexpr int kStageWords = 16;

struct MfmaArgs {
    const uint32_t* pk;   // packed records, [rows*cols + 1][n][re, im]
    const uint32_t* wts;
    int rows, cols, n, Nx, Ny, nul, min_nb;
    int row0;             // first output row of this chunk
    int chunk_rows;
    int TH, TW;           // output pixels per workgroup
    float* gamma;         // [slots][n][n] complex float, interleaved
    int* npix;            // [slots]
    int list_cap;         // ints per wave list
};

// 32x32 accumulator layout: lane l, element v holds C[row][col] with
// col = l % 32 and row = 8*(v/4) + 4*(l/32) + (v%4).
__device__ inline int acc_row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }

template <int NBH> struct TileIdx {
    static constexpr int NB = 2 * NBH;
    // (bi, bj), bi <= bj < NB, in row-major upper order
    static __device__ constexpr int idx(int bi, int bj) {
        return bi * NB - (bi * (bi - 1)) / 2 + (bj - bi);
    }
    static constexpr int count = NB * (NB + 1) / 2;
};

__device__ inline void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool F16> struct Frag;
template <> struct Frag<true> {
    typedef f16x8 type;
    static __device__ inline f32x16 mfma(type a, type b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};
template <> struct Frag<false> {
    typedef bf16x8 type;
    static __device__ inline f32x16 mfma(type a, type b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

template <int NBH, int WAVES, bool F16>
__global__ __launch_bounds__(64 * WAVES) void gram_mfma(MfmaArgs a) {
    typedef typename Frag<F16>::type frag_t;
    constexpr int NB = 2 * NBH;
    constexpr int NT = TileIdx<NBH>::count;
    extern __shared__ __align__(16) char smem[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t wbytes = (size_t)a.list_cap * 4 + 32 * 33 * 4 + 64 * NBH * 4;
    int* list = (int*)(smem + (size_t)wave * wbytes);
    float* xs = (float*)(list + a.list_cap);   // 32 x 33 transpose scratch
    float* dg = xs + 32 * 33;                  // diagonal of Re G

    const int n = a.n, Nx = a.Nx, Ny = a.Ny;
    const int CS = 2 * n;
    const int npx_zero = a.rows * a.cols;      // index of the all-zero record
    const int R0 = a.row0 + blockIdx.y * a.TH, C0 = blockIdx.x * a.TW;
    const int row_end = a.row0 + a.chunk_rows;
    const int Wx = 2 * Nx + 1, Wn = (2 * Ny + 1) * Wx;
    const int m = lane & 31, half = lane >> 5;

    uint32_t coff[NB];
    bool cval[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int date = 32 * (b % NBH) + m;
        cval[b] = date < n;
        coff[b] = cval[b] ? (uint32_t)(2 * date + (b < NBH ? 0 : 1)) : 0u;
    }

    for (int pid = wave; pid < a.TH * a.TW; pid += WAVES) {
        const int r = R0 + pid / a.TW, c = C0 + pid % a.TW;
        if (r >= row_end || r >= a.rows || c >= a.cols) continue;   // wave-uniform
        const long long pix = (long long)r * a.cols + c;
        const int slot = (r - a.row0) * a.cols + c;
        const uint32_t* w = a.wts + pix * a.nul;
        const bool cen = shp_get(w, shp_flat(0, 0, Nx, Ny));

        // ---- neighbour list ----
        int cnt = 0;
        for (int b0 = 0; b0 < Wn; b0 += 64) {
            const int f = b0 + lane;
            bool take = false;
            int q = 0;
            if (cen && f < Wn) {
                const int rr = r + f / Wx - Ny, cc = c + f % Wx - Nx;
                if (rr >= 0 && rr < a.rows && cc >= 0 && cc < a.cols && shp_get(w, f)) {
                    take = true;
                    q = rr * a.cols + cc;
                }
            }
            const uint64_t bal = __ballot(take);
            if (take) {
                const uint64_t below = (lane == 0) ? 0ull : (bal & ((~0ull) >> (64 - lane)));
                list[cnt + __popcll(below)] = q;
            }
            cnt += __popcll(bal);
        }
        if (!(cen && cnt >= a.min_nb)) {
            if (lane == 0) a.npix[slot] = -1;
            continue;
        }
        const int steps = (cnt + 15) >> 4;
        for (int e = cnt + lane; e < steps * 16; e += 64) list[e] = npx_zero;
        wave_sync_lds();

        // ---- MFMA accumulation ----
        f32x16 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[t][v] = 0.0f;

        for (int s = 0; s < steps; ++s) {
            const int* lp = list + 16 * s + 8 * half;
            uint32_t base[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) base[e] = (uint32_t)lp[e] * (uint32_t)CS;
            frag_t fh[NB], fl[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                uint32_t wv[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) wv[e] = a.pk[base[e] + coff[b]];   // valid address for every lane
#pragma unroll
                for (int e = 0; e < 8; ++e) wv[e] = cval[b] ? wv[e] : 0u;
                u32x4 hv, lv;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    hv[e] = __builtin_amdgcn_perm(wv[2 * e + 1], wv[2 * e], 0x05040100u);
                    lv[e] = __builtin_amdgcn_perm(wv[2 * e + 1], wv[2 * e], 0x07060302u);
                }
                fh[b] = __builtin_bit_cast(frag_t, hv);
                fl[b] = __builtin_bit_cast(frag_t, lv);
            }
#pragma unroll
            for (int bi = 0; bi < NB; ++bi)
#pragma unroll
                for (int bj = bi; bj < NB; ++bj) {
                    f32x16& C = acc[TileIdx<NBH>::idx(bi, bj)];
                    C = Frag<F16>::mfma(fh[bi], fh[bj], C);
                    C = Frag<F16>::mfma(fh[bi], fl[bj], C);
                    C = Frag<F16>::mfma(fl[bi], fh[bj], C);
                }
        }

        // ---- fold P into Gamma ----
#pragma unroll
        for (int I = 0; I < NBH; ++I) {
            const f32x16& A0 = acc[TileIdx<NBH>::idx(I, I)];
            const f32x16& A1 = acc[TileIdx<NBH>::idx(NBH + I, NBH + I)];
#pragma unroll
            for (int v = 0; v < 16; ++v)
                if (acc_row(v, lane) == m) dg[32 * I + m] = A0[v] + A1[v];
        }
        float* gout = a.gamma + (size_t)slot * n * n * 2;
#pragma unroll
        for (int I = 0; I < NBH; ++I)
#pragma unroll
            for (int J = I; J < NBH; ++J) {
                const f32x16& X = acc[TileIdx<NBH>::idx(J, NBH + I)];   // P(re_J, im_I)
                wave_sync_lds();
#pragma unroll
                for (int v = 0; v < 16; ++v) xs[acc_row(v, lane) * 33 + m] = X[v];
                wave_sync_lds();
                const f32x16& RR = acc[TileIdx<NBH>::idx(I, J)];
                const f32x16& II = acc[TileIdx<NBH>::idx(NBH + I, NBH + J)];
                const f32x16& RI = acc[TileIdx<NBH>::idx(I, NBH + J)];
                const int j = 32 * J + m;
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int il = acc_row(v, lane);
                    const int i = 32 * I + il;
                    if (i >= n || j >= n || i > j) continue;
                    float gr = 1.0f, gi = 0.0f;
                    if (i != j) {
                        const float re = RR[v] + II[v];
                        const float im = xs[m * 33 + il] - RI[v];   // P(im_i, re_j) - P(re_i, im_j)
                        const float sc = rsqrtf(dg[i] * dg[j]);
                        gr = re * sc;
                        gi = im * sc;
                    }
                    gout[2 * ((size_t)i * n + j)] = gr;
                    gout[2 * ((size_t)i * n + j) + 1] = gi;
                    if (i != j) {
                        gout[2 * ((size_t)j * n + i)] = gr;
                        gout[2 * ((size_t)j * n + i) + 1] = -gi;
                    }
                }
            }
        if (lane == 0) a.npix[slot] = cnt;
        wave_sync_lds();
    }
}

// 64 < n <= 128: the wave of a pixel walks the date-block pairs (I, J),
// I <= J < nb = ceil(n / 32), diagonal pairs first.  Per pair it keeps the
// four 32 x 32 real products P(re_I, re_J), P(im_I, im_J), P(re_I, im_J),
// P(im_I, re_J) (64 accumulator registers whatever n; 12 MFMAs per 16
// positions) and writes Gamma block (I, J) and its mirror; Re G_ii of the
// diagonal pairs (per-wave LDS) normalises the later ones.  Each position's
// (re, im) word pair of a date is one 8-byte load, and records are addressed
// with 64-bit offsets (2n words per record exceed 32-bit word offsets on
// large tiles).
template <int WAVES, bool F16>
__global__ __launch_bounds__(64 * WAVES) void gram_mfma_pairs(MfmaArgs a) {
    typedef typename Frag<F16>::type frag_t;
    extern __shared__ __align__(16) char smem[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t wbytes = (size_t)a.list_cap * 4 + 128 * 4;
    int* list = (int*)(smem + (size_t)wave * wbytes);
    float* dg = (float*)(list + a.list_cap);   // Re G_ii of the pixel, 128 dates

    const int n = a.n, Nx = a.Nx, Ny = a.Ny;
    const int nb = (n + 31) >> 5;
    const int npairs = nb * (nb + 1) / 2;
    const size_t CS = 2 * (size_t)n;
    const int npx_zero = a.rows * a.cols;      // index of the all-zero record
    const int R0 = a.row0 + blockIdx.y * a.TH, C0 = blockIdx.x * a.TW;
    const int row_end = a.row0 + a.chunk_rows;
    const int Wx = 2 * Nx + 1, Wn = (2 * Ny + 1) * Wx;
    const int m = lane & 31, half = lane >> 5;

    for (int pid = wave; pid < a.TH * a.TW; pid += WAVES) {
        const int r = R0 + pid / a.TW, c = C0 + pid % a.TW;
        if (r >= row_end || r >= a.rows || c >= a.cols) continue;   // wave-uniform
        const long long pix = (long long)r * a.cols + c;
        const int slot = (r - a.row0) * a.cols + c;
        const uint32_t* w = a.wts + pix * a.nul;
        const bool cen = shp_get(w, shp_flat(0, 0, Nx, Ny));

        // ---- neighbour list ----
        int cnt = 0;
        for (int b0 = 0; b0 < Wn; b0 += 64) {
            const int f = b0 + lane;
            bool take = false;
            int q = 0;
            if (cen && f < Wn) {
                const int rr = r + f / Wx - Ny, cc = c + f % Wx - Nx;
                if (rr >= 0 && rr < a.rows && cc >= 0 && cc < a.cols && shp_get(w, f)) {
                    take = true;
                    q = rr * a.cols + cc;
                }
            }
            const uint64_t bal = __ballot(take);
            if (take) {
                const uint64_t below = (lane == 0) ? 0ull : (bal & ((~0ull) >> (64 - lane)));
                list[cnt + __popcll(below)] = q;
            }
            cnt += __popcll(bal);
        }
        if (!(cen && cnt >= a.min_nb)) {
            if (lane == 0) a.npix[slot] = -1;
            continue;
        }
        const int steps = (cnt + 15) >> 4;
        for (int e = cnt + lane; e < steps * 16; e += 64) list[e] = npx_zero;
        wave_sync_lds();

        float2* gout = (float2*)a.gamma + (size_t)slot * n * n;
        for (int p = 0; p < npairs; ++p) {   // uniform
            int I = p, J = p;
            if (p >= nb) {   // off-diagonal pairs in row order
                int q = p - nb;
                I = 0;
                while (q >= nb - 1 - I) {
                    q -= nb - 1 - I;
                    ++I;
                }
                J = I + 1 + q;
            }
            const int di = 32 * I + m, dj = 32 * J + m;
            const bool vi = di < n, vj = dj < n;
            const int oi = vi ? di : 0, oj = vj ? dj : 0;   // in (re, im) word pairs
            f32x16 RR, II, RI, IR;
#pragma unroll
            for (int v = 0; v < 16; ++v) RR[v] = II[v] = RI[v] = IR[v] = 0.0f;
            for (int s = 0; s < steps; ++s) {
                const int* lp = list + 16 * s + 8 * half;
                uint2 wi[8], wj[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const uint2* rec = (const uint2*)(a.pk + (size_t)(uint32_t)lp[e] * CS);
                    wi[e] = rec[oi];   // valid address for every lane
                    wj[e] = rec[oj];
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    if (!vi) wi[e] = make_uint2(0u, 0u);
                    if (!vj) wj[e] = make_uint2(0u, 0u);
                }
                u32x4 rih, ril, iih, iil, rjh, rjl, ijh, ijl;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    rih[e] = __builtin_amdgcn_perm(wi[2 * e + 1].x, wi[2 * e].x, 0x05040100u);
                    ril[e] = __builtin_amdgcn_perm(wi[2 * e + 1].x, wi[2 * e].x, 0x07060302u);
                    iih[e] = __builtin_amdgcn_perm(wi[2 * e + 1].y, wi[2 * e].y, 0x05040100u);
                    iil[e] = __builtin_amdgcn_perm(wi[2 * e + 1].y, wi[2 * e].y, 0x07060302u);
                    rjh[e] = __builtin_amdgcn_perm(wj[2 * e + 1].x, wj[2 * e].x, 0x05040100u);
                    rjl[e] = __builtin_amdgcn_perm(wj[2 * e + 1].x, wj[2 * e].x, 0x07060302u);
                    ijh[e] = __builtin_amdgcn_perm(wj[2 * e + 1].y, wj[2 * e].y, 0x05040100u);
                    ijl[e] = __builtin_amdgcn_perm(wj[2 * e + 1].y, wj[2 * e].y, 0x07060302u);
                }
                const frag_t Rh = __builtin_bit_cast(frag_t, rih), Rl = __builtin_bit_cast(frag_t, ril);
                const frag_t Ih = __builtin_bit_cast(frag_t, iih), Il = __builtin_bit_cast(frag_t, iil);
                const frag_t Sh = __builtin_bit_cast(frag_t, rjh), Sl = __builtin_bit_cast(frag_t, rjl);
                const frag_t Th = __builtin_bit_cast(frag_t, ijh), Tl = __builtin_bit_cast(frag_t, ijl);
                RR = Frag<F16>::mfma(Rh, Sh, RR);
                RR = Frag<F16>::mfma(Rh, Sl, RR);
                RR = Frag<F16>::mfma(Rl, Sh, RR);
                II = Frag<F16>::mfma(Ih, Th, II);
                II = Frag<F16>::mfma(Ih, Tl, II);
                II = Frag<F16>::mfma(Il, Th, II);
                RI = Frag<F16>::mfma(Rh, Th, RI);
                RI = Frag<F16>::mfma(Rh, Tl, RI);
                RI = Frag<F16>::mfma(Rl, Th, RI);
                IR = Frag<F16>::mfma(Ih, Sh, IR);
                IR = Frag<F16>::mfma(Ih, Sl, IR);
                IR = Frag<F16>::mfma(Il, Sh, IR);
            }
            if (I == J) {
#pragma unroll
                for (int v = 0; v < 16; ++v)
                    if (acc_row(v, lane) == m) dg[32 * I + m] = RR[v] + II[v];
                wave_sync_lds();
            }
            const int j = 32 * J + m;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int i = 32 * I + acc_row(v, lane);
                if (i >= n || j >= n || i > j) continue;
                float gr = 1.0f, gi = 0.0f;
                if (i != j) {
                    const float sc = rsqrtf(dg[i] * dg[j]);
                    gr = (RR[v] + II[v]) * sc;
                    gi = (IR[v] - RI[v]) * sc;   // P(im_i, re_j) - P(re_i, im_j)
                }
                gout[(size_t)i * n + j] = make_float2(gr, gi);
                if (i != j) gout[(size_t)j * n + i] = make_float2(gr, -gi);
            }
        }
        if (lane == 0) a.npix[slot] = cnt;
        wave_sync_lds();
    }
}

// ---------------------------------------------------------------------------
// n <= 32: the records of the union of a workgroup's windows are staged in
// LDS once, and the waves take their pixels' neighbour lists from there.
// A TH x TW pixel tile needs (TH + 2Ny) x (TW + 2Nx) positions: for 4 x 8
// pixels and a 15 x 11 window that is 308 records for 32 pixels, against
// 165 per pixel gathered one by one.  All four real 32 x 32 tiles
// P(re,re), P(im,im), P(re,im), P(im,re) are accumulated, so the fold needs
// no transpose: Re G = RR + II, Im G(i,j) = P(im_i,re_j) - P(re_i,im_j).
// Workgroups are numbered so that each XCD (which owns an L2) takes a
// contiguous run of tiles and the halos neighbouring tiles share stay in it.
// ---------------------------------------------------------------------------
template <bool F16, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void gram_mfma_lds(MfmaArgs a) {
    constexpr int NT = 64 * WAVES;
    typedef typename Frag<F16>::type frag_t;
    extern __shared__ __align__(16) char smem[];
    const int n = a.n, Nx = a.Nx, Ny = a.Ny;
    const int TH = a.TH, TW = a.TW;
    const int WR = TH + 2 * Ny, WC = TW + 2 * Nx, NPOS = WR * WC;
    uint2* rec = (uint2*)smem;                             // [NPOS + 1][32] (re, im) words
    // wave index in a scalar register, so the pixel loop below is uniform
    // control flow (scalar branches) rather than exec-masked
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    int* list = (int*)(rec + (size_t)(NPOS + 1) * 32) + wave * (a.list_cap + 32 + 64);
    float* dg = (float*)(list + a.list_cap);                // diagonal of Re G, 32
    uint32_t* wbuf = (uint32_t*)(dg + 32);                  // next pixel's mask words, 64

    // XCD-aware tile order (blocks are dealt to the 8 XCDs round-robin)
    const unsigned nb = gridDim.x * gridDim.y, bid = blockIdx.y * gridDim.x + blockIdx.x;
    const unsigned xcd = bid % 8u, q = nb / 8u, rem = nb % 8u;
    const unsigned logical = xcd * q + min(xcd, rem) + bid / 8u;
    const int tx = (int)(logical % gridDim.x), ty = (int)(logical / gridDim.x);
    const int R0 = a.row0 + ty * TH, C0 = tx * TW;
    const int row_end = a.row0 + a.chunk_rows;
    const int rbase = R0 - Ny, cbase = C0 - Nx;

    // ---- stage the records, padded to 32 dates (zeros beyond n and outside
    //      the image), eight loads in flight ----
    {
        const int total = (NPOS + 1) * 32;
        const uint2* src = (const uint2*)a.pk;
        for (int e0 = threadIdx.x; e0 < total; e0 += 8 * NT) {
            uint2 v[8];
            bool in[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + u * NT;
                const int pos = e >> 5, d = e & 31;
                const int rr = rbase + pos / WC, cc = cbase + pos % WC;
                in[u] = e < total && d < n && pos < NPOS && rr >= 0 && rr < a.rows && cc >= 0 && cc < a.cols;
                // slots without data read record 0 and are zeroed below
                const size_t at = in[u] ? ((size_t)rr * a.cols + cc) * n + d : 0;
                v[u] = src[at];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + u * NT;
                if (e < total) rec[e] = in[u] ? v[u] : make_uint2(0u, 0u);
            }
        }
    }
    __syncthreads();

    const int Wx = 2 * Nx + 1, Wn = (2 * Ny + 1) * Wx;
    const int m = lane & 31, hf = lane >> 5;
    // window offsets of this lane in every 64-wide pass (pixel-independent)
    constexpr int kPass = (kStageWords * 32 + 63) / 64;
    int offy[kPass], offx[kPass];
#pragma unroll
    for (int it = 0; it < kPass; ++it) {
        const int f = it * 64 + lane;
        offy[it] = f / Wx;
        offx[it] = f - offy[it] * Wx;
    }

    // mask words of a pixel in lanes 0..nul-1 (one load).  The next pixel's
    // are requested at the top of an iteration and parked in LDS before this
    // pixel's Gamma stores, so no wait for them ever waits for the stores.
    auto mask_word = [&](int pid) -> uint32_t {
        const int r = R0 + pid / TW, c = C0 + pid % TW;
        const bool ok = pid < TH * TW && r < row_end && r < a.rows && c < a.cols;
        // always one load, of pixel 0 when there is no pixel (such a pixel is
        // skipped before its word is used)
        const long long px = ok ? (long long)r * a.cols + c : 0;
        return a.wts[px * a.nul + (lane < a.nul ? lane : 0)];
    };
    wbuf[lane] = mask_word(wave);
    for (int pid = wave; pid < TH * TW; pid += WAVES) {
        const uint32_t wword = wbuf[lane];
        const uint32_t wnext = mask_word(pid + WAVES);
        const int r = R0 + pid / TW, c = C0 + pid % TW;
        if (r >= row_end || r >= a.rows || c >= a.cols) {   // wave-uniform
            wbuf[lane] = wnext;
            continue;
        }
        const long long pix = (long long)r * a.cols + c;
        const int slot = (r - a.row0) * a.cols + c;
        const int fc = shp_flat(0, 0, Nx, Ny);
        // mask words reach the lanes through readlane (scalar registers), not
        // through an LDS-routed shuffle
        const bool cen = (__builtin_amdgcn_readlane(wword, fc >> 5) >> (fc & 31)) & 1u;
        // ---- neighbour list (staged record indices) ----
        const int lr = r - R0, lc = c - C0;                  // window origin in the stage
        int cnt = 0;
#pragma unroll
        for (int it = 0; it < kPass; ++it) {
            if (it * 64 >= Wn) break;   // uniform
            const int f = it * 64 + lane;
            // lanes 0..31 test word 2it, lanes 32..63 word 2it+1 (lanes past
            // the last word hold word 0 and only cover positions >= Wn)
            const uint32_t wlo = __builtin_amdgcn_readlane(wword, 2 * it);
            const uint32_t whi = __builtin_amdgcn_readlane(wword, 2 * it + 1);
            const uint32_t wf = lane < 32 ? wlo : whi;
            bool take = false;
            int qpos = 0;
            if (cen && f < Wn) {
                const int fy = offy[it], fx = offx[it];
                const int rr = r + fy - Ny, cc = c + fx - Nx;
                if (rr >= 0 && rr < a.rows && cc >= 0 && cc < a.cols && ((wf >> (f & 31)) & 1u)) {
                    take = true;
                    qpos = (lr + fy) * WC + (lc + fx);
                }
            }
            const uint64_t bal = __ballot(take);
            if (take) {
                const uint64_t below = (lane == 0) ? 0ull : (bal & ((~0ull) >> (64 - lane)));
                list[cnt + __popcll(below)] = qpos;
            }
            cnt += __popcll(bal);
        }
        if (!(cen && cnt >= a.min_nb)) {
            wbuf[lane] = wnext;
            if (lane == 0) a.npix[slot] = -1;
            continue;
        }
        const int steps = (cnt + 15) >> 4;
        for (int e = cnt + lane; e < steps * 16; e += 64) list[e] = NPOS;
        wave_sync_lds();

        f32x16 RR, II, RI, IR;
#pragma unroll
        for (int v = 0; v < 16; ++v) RR[v] = II[v] = RI[v] = IR[v] = 0.0f;
        // operands of step st + 1 are loaded while step st's MFMAs run; one
        // 8-byte read gives this lane's (re, im) words of a position
        uint2 x[8];
        auto load_step = [&](int st, uint2 (&y)[8]) {
            const int* lp = list + 16 * st + 8 * hf;
#pragma unroll
            for (int e = 0; e < 8; ++e) y[e] = rec[(uint32_t)lp[e] * 32u + (uint32_t)m];
        };
        load_step(0, x);
        for (int st = 0; st < steps; ++st) {
            uint2 nx[8];
            load_step(st + 1 < steps ? st + 1 : st, nx);
            u32x4 hr, lr4, hi4, li4;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                hr[e] = __builtin_amdgcn_perm(x[2 * e + 1].x, x[2 * e].x, 0x05040100u);
                lr4[e] = __builtin_amdgcn_perm(x[2 * e + 1].x, x[2 * e].x, 0x07060302u);
                hi4[e] = __builtin_amdgcn_perm(x[2 * e + 1].y, x[2 * e].y, 0x05040100u);
                li4[e] = __builtin_amdgcn_perm(x[2 * e + 1].y, x[2 * e].y, 0x07060302u);
            }
            const frag_t fhr = __builtin_bit_cast(frag_t, hr), flr = __builtin_bit_cast(frag_t, lr4);
            const frag_t fhi = __builtin_bit_cast(frag_t, hi4), fli = __builtin_bit_cast(frag_t, li4);
            // interleaved so that consecutive MFMAs feed different accumulators
            RR = Frag<F16>::mfma(fhr, fhr, RR);
            II = Frag<F16>::mfma(fhi, fhi, II);
            RI = Frag<F16>::mfma(fhr, fhi, RI);
            IR = Frag<F16>::mfma(fhi, fhr, IR);
            RR = Frag<F16>::mfma(fhr, flr, RR);
            II = Frag<F16>::mfma(fhi, fli, II);
            RI = Frag<F16>::mfma(fhr, fli, RI);
            IR = Frag<F16>::mfma(fhi, flr, IR);
            RR = Frag<F16>::mfma(flr, fhr, RR);
            II = Frag<F16>::mfma(fli, fhi, II);
            RI = Frag<F16>::mfma(flr, fhi, RI);
            IR = Frag<F16>::mfma(fli, fhr, IR);
#pragma unroll
            for (int e = 0; e < 8; ++e) x[e] = nx[e];
        }

        wbuf[lane] = wnext;   // waits for the load here, before the stores below

        // ---- fold P into Gamma ----
        // Re G_mm sits in register 4 (m / 8) + m % 4 of the lanes whose half
        // holds row m ((m / 4) % 2 == hf); it is picked out by selects, and
        // every later LDS read is issued before any result is used.
        float dv = 0.0f;
#pragma unroll
        for (int v = 0; v < 16; ++v) dv = (acc_row(v, lane) == m) ? RR[v] + II[v] : dv;
        // published as 1/sqrt(Re G_mm): one transcendental per lane, and the
        // normalisation of entry (i, m) is a product
        if (((m >> 2) & 1) == hf) dg[m] = rsqrtf(dv);
        wave_sync_lds();
        float di[16];
#pragma unroll
        for (int v = 0; v < 16; ++v) di[v] = dg[acc_row(v, lane)];
        const float dj = dg[m];
        float2* gout = reinterpret_cast<float2*>(a.gamma) + (size_t)slot * n * n;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int i = acc_row(v, lane);
            const float sc = di[v] * dj;
            const bool off = i != m;
            const float gr = off ? (RR[v] + II[v]) * sc : 1.0f;
            const float gi = off ? (IR[v] - RI[v]) * sc : 0.0f;   // P(im_i, re_j) - P(re_i, im_j)
            if (i < n && m < n) gout[(size_t)i * n + m] = make_float2(gr, gi);
        }
        if (lane == 0) a.npix[slot] = cnt;
        wave_sync_lds();
    }
}

static size_t gram_lds_bytes(int Nx, int Ny, int TH, int TW, int list_cap, int waves) {
    // records are staged padded to 32 dates
    const size_t stage = (size_t)((TH + 2 * Ny) * (TW + 2 * Nx) + 1) * 32 * 8;
    return stage + (size_t)waves * (list_cap + 32 + 64) * 4;
}

static int gram_list_cap(int Nx, int Ny) {
    const int Wn = (2 * Ny + 1) * (2 * Nx + 1);
    return ((Wn + 15) / 16) * 16 + 16;
}

// Tile of the staged kernel.  First choice: an 8 x 16 tile for twelve
// waves (three per SIMD), one workgroup per CU, which stages ~4.2 records
// per pixel at a 15 x 11 window; then a 4 x 16 tile for eight waves (~6.6;
// a 3 x 8 tile stages ~11.9).  Otherwise the largest TH x TW tile
// (TH <= 4, TW <= 8) for four waves whose stage keeps two workgroups on a
// CU.  Returns the stage's LDS bytes.
// `only`: a scheduled variant (FRINGE_GRAM_STAGED12 / 8 / 4) tries that tile
// alone; FRINGE_GRAM_DEFAULT tries them in order.
static size_t gram_stage_tile(int Nx, int Ny, int& TH, int& TW, int& waves, int only = FRINGE_GRAM_DEFAULT) {
    const int cap = gram_list_cap(Nx, Ny);
    waves = 12;
    TH = 8;
    TW = 16;
    if ((only == FRINGE_GRAM_DEFAULT || only == FRINGE_GRAM_STAGED12) &&
        gram_lds_bytes(Nx, Ny, TH, TW, cap, waves) <= 156 * 1024)
        return gram_lds_bytes(Nx, Ny, TH, TW, cap, waves);
    waves = 8;
    TH = 4;
    TW = 16;
    if ((only == FRINGE_GRAM_DEFAULT || only == FRINGE_GRAM_STAGED8) &&
        gram_lds_bytes(Nx, Ny, TH, TW, cap, waves) <= 150 * 1024)
        return gram_lds_bytes(Nx, Ny, TH, TW, cap, waves);
    if (only != FRINGE_GRAM_DEFAULT && only != FRINGE_GRAM_STAGED4) return (size_t)1 << 40;
    const size_t budget = 78 * 1024;
    waves = 4;
    TH = 4;
    TW = 8;
    while (TH > 1 && gram_lds_bytes(Nx, Ny, TH, TW, cap, waves) > budget) --TH;
    while (TW > 1 && gram_lds_bytes(Nx, Ny, TH, TW, cap, waves) > budget) TW >>= 1;
    if (TH * TW >= 4) return gram_lds_bytes(Nx, Ny, TH, TW, cap, waves);
    // a stage for fewer than four pixels leaves waves idle: take the largest
    // tile of at least four pixels that fits one workgroup per CU, or none
    // (the unstaged kernel then gathers from L2)
    for (TH = 4; TH >= 1; --TH)
        for (TW = 8; TW >= 1; TW >>= 1)
            if (TH * TW >= 4 && gram_lds_bytes(Nx, Ny, TH, TW, cap, waves) <= 156 * 1024)
                return gram_lds_bytes(Nx, Ny, TH, TW, cap, waves);
    return (size_t)1 << 40;
}

// The staged kernel takes windows of up to kStageWords mask words; larger
// windows go to the unstaged kernel, which walks any number of words.
// A scheduled variant (fringe_schedule_set) that does not fit the window
// falls back to the default choice.
static bool gram_staged(int n, int Nx, int Ny, int& TH, int& TW, int& waves, size_t& bytes) {
    if (n > 32 || shp_words(Nx, Ny) > kStageWords) return false;
    const int v = schedule_lookup(FRINGE_SCHED_GRAM_VARIANT, n, Nx, Ny, FRINGE_GRAM_DEFAULT);
    if (v == FRINGE_GRAM_UNSTAGED) return false;
    if (v != FRINGE_GRAM_DEFAULT) {
        bytes = gram_stage_tile(Nx, Ny, TH, TW, waves, v);
        if (bytes <= 160 * 1024) return true;
    }
    bytes = gram_stage_tile(Nx, Ny, TH, TW, waves);
    return bytes <= 160 * 1024;
}

int gram_stage_rows(int n, int Nx, int Ny) {
    int TH, TW, waves;
    size_t bytes;
    return gram_staged(n, Nx, Ny, TH, TW, waves, bytes) ? TH : 1;
}

// packed records followed by the per-date maxima (256-byte aligned)
size_t mfma_pack_bytes(int rows, int cols, int n) {
    const size_t rec = ((size_t)rows * cols + 1) * 2 * n * sizeof(uint32_t);
    return ((rec + 255) & ~(size_t)255) + (size_t)n * sizeof(uint32_t);
}

int pack_launch(const float2* slc, int rows, int cols, int n, bool f16, uint32_t* pk, hipStream_t s) {
    const long long npx = (long long)rows * cols;
    const size_t rec = ((size_t)npx + 1) * 2 * n * sizeof(uint32_t);
    uint32_t* dmax = (uint32_t*)((char*)pk + ((rec + 255) & ~(size_t)255));
    const unsigned blocks = (unsigned)((npx + 63) / 64);
    if (f16) {
        hipLaunchKernelGGL(date_absmax_clear, dim3((n + 63) / 64), dim3(64), 0, s, dmax, n);
        const unsigned bx = (unsigned)std::min<long long>((npx + 255) / 256, 64);   // 64 workgroups per date
        hipLaunchKernelGGL(date_absmax, dim3(bx, n), dim3(256), 0, s, slc, npx, dmax);
        hipLaunchKernelGGL(pack_hilo<true>, dim3(blocks), dim3(256), (size_t)64 * 2 * n * 4, s, slc, n,
                           npx, (const uint32_t*)dmax, pk);
    } else {
        hipLaunchKernelGGL(pack_hilo<false>, dim3(blocks), dim3(256), (size_t)64 * 2 * n * 4, s, slc, n,
                           npx, (const uint32_t*)dmax, pk);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int NBH, int WAVES, bool F16>
static void launch_mfma(MfmaArgs a, hipStream_t s) {
    const size_t wbytes = (size_t)a.list_cap * 4 + 32 * 33 * 4 + 64 * NBH * 4;
    const dim3 grid((unsigned)((a.cols + a.TW - 1) / a.TW),
                    (unsigned)((a.chunk_rows + a.TH - 1) / a.TH));
    hipLaunchKernelGGL((gram_mfma<NBH, WAVES, F16>), grid, dim3(64 * WAVES), wbytes * WAVES, s, a);
}

int gram_mfma_launch(const LinkPlan& P, const uint32_t* pk, const uint32_t* wts, long long first,
                     int slots, hipStream_t s) {
    if (P.n > kMaxBandsMfma || first % P.cols != 0 || slots % P.cols != 0) return -1;
    MfmaArgs a{};
    a.pk = pk;
    a.wts = wts;
    a.rows = P.rows;
    a.cols = P.cols;
    a.n = P.n;
    a.Nx = P.Nx;
    a.Ny = P.Ny;
    a.nul = P.nul;
    a.min_nb = P.min_nb;
    a.row0 = (int)(first / P.cols);
    a.chunk_rows = slots / P.cols;
    a.TH = 2;
    a.TW = 8;
    a.gamma = (float*)P.d_gamma;
    a.npix = P.d_npix;
    a.list_cap = gram_list_cap(P.Nx, P.Ny);
    const bool f16 = (P.precision != FRINGE_PREC_MIXED_BF16);
    if (P.n <= 32) {
        int TH, TW, waves;
        size_t bytes = 0;
        if (gram_staged(P.n, P.Nx, P.Ny, TH, TW, waves, bytes)) {
            a.TH = TH;
            a.TW = TW;
            const dim3 grid((unsigned)((a.cols + TW - 1) / TW), (unsigned)((a.chunk_rows + TH - 1) / TH));
            const dim3 blk(64 * waves);
            if (waves == 12) {
                if (f16) hipLaunchKernelGGL((gram_mfma_lds<true, 12>), grid, blk, bytes, s, a);
                else hipLaunchKernelGGL((gram_mfma_lds<false, 12>), grid, blk, bytes, s, a);
            } else if (waves == 8) {
                if (f16) hipLaunchKernelGGL((gram_mfma_lds<true, 8>), grid, blk, bytes, s, a);
                else hipLaunchKernelGGL((gram_mfma_lds<false, 8>), grid, blk, bytes, s, a);
            } else {
                if (f16) hipLaunchKernelGGL((gram_mfma_lds<true, 4>), grid, blk, bytes, s, a);
                else hipLaunchKernelGGL((gram_mfma_lds<false, 4>), grid, blk, bytes, s, a);
            }
        } else if (f16) {
            launch_mfma<1, 4, true>(a, s);
        } else {
            launch_mfma<1, 4, false>(a, s);
        }
    } else {   // above 32 dates
        const size_t wbytes = (size_t)a.list_cap * 4 + 128 * 4;
        const dim3 grid((unsigned)((a.cols + a.TW - 1) / a.TW), (unsigned)((a.chunk_rows + a.TH - 1) / a.TH));
        if (f16) hipLaunchKernelGGL((gram_mfma_pairs<4, true>), grid, dim3(256), wbytes * 4, s, a);
        else hipLaunchKernelGGL((gram_mfma_pairs<4, false>), grid, dim3(256), wbytes * 4, s, a);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace fringe
