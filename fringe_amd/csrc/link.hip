// link.hip -- per-pixel phase linking on gfx950: masked-neighbourhood
// coherence (FP64 "plain" Gram) and the per-pixel estimator chain.
//
// Reference semantics: src/evd/evd.cpp:512-788 (EVD / MLE / STBAS) and
// src/phase_link/phase_link.cpp:479-666 (EMI with EVD fallback).  The LAPACK
// calls of EigenLapack.hpp (zheevr RANGE='I', zpotrf + zpotri) are replaced by
// a wave-resident Hermitian solver:
//   * Householder tridiagonalisation A = Q T Q^H held in LDS, one lane per row;
//   * the wanted eigenvalue of T by multisection (every lane of the group
//     evaluates one Sturm count per round);
//   * its eigenvector by inverse iteration on T (pivoted tridiagonal solve);
//   * back-transformation by the stored reflectors;
//   * the inverse of the real |Gamma| by a Gauss-Jordan sweep in FP64, whose
//     pivots are the Cholesky pivots, so "not positive definite" is decided
//     the way zpotrf decides it.
#include "device_common.hpp"
#include "engine_internal.hpp"
#include "link_solve.hpp"
#include "solve32.hpp"
#include "timing.hpp"

namespace fringe {

// ---------------------------------------------------------------------------
// FP64 Gram: one group of G lanes per pixel, lane g owns rows g + G*q (q < R)
// and accumulates them against a chunk of CB columns.  The masked,
// in-image window positions are compacted into a list first, so the work is
// proportional to popcount(mask), not to the window.
// ---------------------------------------------------------------------------
struct GramArgs {
    const float2* slc;
    const uint32_t* wts;
    int rows, cols, n, Nx, Ny, nul, min_nb;
    long long npx;
    long long first;   // linear pixel index (row*cols + col) of slot 0
    int slots;
    size_t group_bytes;
    void* gamma;       // [slots][n][n] cx<double>
    int* npix;         // [slots]
};

template <int G>
__device__ inline int build_position_list(const uint32_t* w, bool cen, int r0, int c0, int rows,
                                          int cols, int Nx, int Ny, int* list) {
    const int g = threadIdx.x % G;
    const int Wx = 2 * Nx + 1, Wn = (2 * Ny + 1) * Wx;
    int cnt = 0;
    for (int b0 = 0; b0 < Wn; b0 += G) {
        const int f = b0 + g;
        bool take = false;
        int q = 0;
        if (cen && f < Wn) {
            const int rr = r0 + f / Wx - Ny, cc = c0 + f % Wx - Nx;
            if (rr >= 0 && rr < rows && cc >= 0 && cc < cols && shp_get(w, f)) {
                take = true;
                q = rr * cols + cc;
            }
        }
        const uint64_t bal = group_ballot<G>(take);
        if (take) {
            const uint64_t below = (g == 0) ? 0ull : (bal & ((~0ull) >> (64 - g)));
            list[cnt + __popcll(below)] = q;
        }
        cnt += __popcll(bal);
    }
    return cnt;
}

template <int G, int R, int CB>
__global__ __launch_bounds__(64) void gram_f64(GramArgs a) {
    constexpr int GPW = 64 / G;
    const int grp = threadIdx.x / G, g = threadIdx.x % G;
    const int slot = blockIdx.x * GPW + grp;
    const int n = a.n;
    const int Wn = (2 * a.Ny + 1) * (2 * a.Nx + 1);
    const int NC = (n + CB - 1) / CB;
    extern __shared__ __align__(16) char smem[];
    char* base = smem + grp * a.group_bytes;
    int* list = (int*)base;
    cx<double>* zb = (cx<double>*)(base + ((Wn * 4 + 15) & ~15));
    double* dg = (double*)(zb + NC * CB);

    const bool inslot = slot < a.slots;
    const long long pix = a.first + (inslot ? slot : 0);
    const int r0 = (int)(pix / a.cols), c0 = (int)(pix % a.cols);
    const uint32_t* w = a.wts + pix * a.nul;
    const bool cen = inslot && shp_get(w, shp_flat(0, 0, a.Nx, a.Ny));
    const int cnt = build_position_list<G>(w, cen, r0, c0, a.rows, a.cols, a.Nx, a.Ny, list);
    const bool go = cen && cnt >= a.min_nb;
    const int K = go ? cnt : 0;
    int Kw = K;
    if (GPW > 1) Kw = max(Kw, __shfl_xor(Kw, 32));
    for (int c = g; c < NC * CB; c += G) zb[c] = cx<double>{0.0, 0.0};
    __syncthreads();

    cx<double>* gam = (cx<double>*)a.gamma + (size_t)(inslot ? slot : 0) * n * n;
    for (int cb = 0; cb < NC; ++cb) {
        cx<double> acc[R][CB];
        double d[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            d[q] = 0.0;
#pragma unroll
            for (int c = 0; c < CB; ++c) acc[q][c] = cx<double>{0.0, 0.0};
        }
        for (int k = 0; k < Kw; ++k) {
            const int pos = (k < K) ? list[k] : -1;
            cx<double> z[R];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int r = g + G * q;
                z[q] = cx<double>{0.0, 0.0};
                if (pos >= 0 && r < n) {
                    const float2 v = a.slc[(long long)r * a.npx + pos];
                    z[q] = cx<double>{(double)v.x, (double)v.y};
                }
                if (r < n) zb[r] = z[q];
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < R; ++q) {
#pragma unroll
                for (int c = 0; c < CB; ++c) cfma_c(acc[q][c], z[q], zb[cb * CB + c]);
                if (cb == 0) d[q] += cabs2(z[q]);
            }
            __syncthreads();
        }
        if (cb == 0) {
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int r = g + G * q;
                if (r < n) dg[r] = d[q];
            }
            __syncthreads();
        }
        if (go) {
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int r = g + G * q;
                if (r >= n) continue;
#pragma unroll
                for (int c = 0; c < CB; ++c) {
                    const int col = cb * CB + c;
                    if (col >= n || col < r) continue;
                    if (col == r) {
                        gam[(size_t)r * n + r] = cx<double>{1.0, 0.0};
                    } else {
                        // Gamma_ij = sum z_i conj(z_j) / sqrt(D_i D_j) (evd.cpp:573-579)
                        const double s = 1.0 / sqrt(dg[r] * dg[col]);
                        const cx<double> v = cscale(acc[q][c], s);
                        gam[(size_t)r * n + col] = v;
                        gam[(size_t)col * n + r] = cconj(v);
                    }
                }
            }
        }
        __syncthreads();
    }
    if (inslot && g == 0) a.npix[slot] = go ? cnt : -1;
}

static size_t gram_f64_group_bytes(int n, int Nx, int Ny, int CB) {
    const int Wn = (2 * Ny + 1) * (2 * Nx + 1);
    const int NC = (n + CB - 1) / CB;
    size_t b = ((size_t)Wn * 4 + 15) & ~(size_t)15;
    b += (size_t)NC * CB * sizeof(cx<double>);
    b += (size_t)n * sizeof(double);
    return (b + 15) & ~(size_t)15;
}

template <int G, int R, int CB>
static void launch_gram_f64(GramArgs a, hipStream_t s) {
    constexpr int GPW = 64 / G;
    a.group_bytes = gram_f64_group_bytes(a.n, a.Nx, a.Ny, CB);
    const unsigned blocks = (unsigned)((a.slots + GPW - 1) / GPW);
    hipLaunchKernelGGL((gram_f64<G, R, CB>), dim3(blocks), dim3(64), a.group_bytes * GPW, s, a);
}

// ---------------------------------------------------------------------------
// Solver launch
// ---------------------------------------------------------------------------
template <typename T, typename Tg, int G, int R>
static void launch_solve(SolveArgs a, hipStream_t s) {
    constexpr int GPW = G >= 64 ? 1 : 64 / G;
    a.group_bytes = solve_group_bytes<T>(a.n);
    const unsigned blocks = (unsigned)((a.slots + GPW - 1) / GPW);
    hipLaunchKernelGGL((solve_kernel<T, Tg, G, R>), dim3(blocks), dim3(G > 64 ? G : 64), a.group_bytes * GPW, s,
                       a);
}

// Static LDS of the solver kernel that runs above 64 dates (its reduction
// scratch); the launch check adds it to the dynamic LDS of a pixel.
size_t solve_static_lds(int precision) {
    hipFuncAttributes fa{};
    const void* k = precision == FRINGE_PREC_FP64 ? (const void*)solve_kernel<double, double, 128, 1>
                                                   : (const void*)solve_kernel<float, double, 128, 1>;
    return hipFuncGetAttributes(&fa, k) == hipSuccess ? fa.sharedSizeBytes : 256;
}

template <typename T, typename Tg>
static void dispatch_solve(const SolveArgs& a, hipStream_t s) {
    if (a.n <= 32) {
        const unsigned blocks = (unsigned)((a.slots + 1) / 2);
        if (a.module == FRINGE_MODULE_PHASE_LINK)
            hipLaunchKernelGGL((solve_reg_kernel<T, Tg, MODE_EMI, 32>), dim3(blocks), dim3(64), 0, s, a);
        else if (a.method == FRINGE_METHOD_MLE)
            hipLaunchKernelGGL((solve_reg_kernel<T, Tg, MODE_MLE, 32>), dim3(blocks), dim3(64), 0, s, a);
        else
            hipLaunchKernelGGL((solve_reg_kernel<T, Tg, MODE_EVD, 32>), dim3(blocks), dim3(64), 0, s, a);
    }
    else if (a.n <= 64) {
        const unsigned blocks = (unsigned)a.slots;
        if (a.module == FRINGE_MODULE_PHASE_LINK)
            hipLaunchKernelGGL((solve_reg_kernel<T, Tg, MODE_EMI, 64>), dim3(blocks), dim3(64), 0, s, a);
        else if (a.method == FRINGE_METHOD_MLE)
            hipLaunchKernelGGL((solve_reg_kernel<T, Tg, MODE_MLE, 64>), dim3(blocks), dim3(64), 0, s, a);
        else
            hipLaunchKernelGGL((solve_reg_kernel<T, Tg, MODE_EVD, 64>), dim3(blocks), dim3(64), 0, s, a);
    }
    else launch_solve<T, Tg, 128, 1>(a, s);   // 64 < n <= 128: two waves per pixel
}

bool link_uses_mfma(int n, int precision) { return precision != FRINGE_PREC_FP64 && n <= 64; }

size_t link_gamma_bytes_per_pixel(int n, int precision) {
    const size_t e = link_uses_mfma(n, precision) ? 8 : 16;
    return (size_t)n * n * e + sizeof(int);
}

// Coherence matrices only (the Gram stage of link_launch), into a caller
// buffer covering the output rows: complex64 from the MFMA Gram, complex128
// from the FP64 one.
int coherence_launch(const LinkPlan& P, const float2* slc, const uint32_t* wts, void* gamma, int* npix,
                     hipStream_t s) {
    const long long first = (long long)P.row_begin * P.cols;
    const int slots = (P.row_end - P.row_begin) * P.cols;
    if (slots <= 0) return 0;
    if (link_uses_mfma(P.n, P.precision)) {
        if (pack_launch(slc, P.rows, P.cols, P.n, P.precision != FRINGE_PREC_MIXED_BF16, P.d_pack, s) != 0)
            return -1;
        LinkPlan Q = P;
        Q.d_gamma = gamma;
        Q.d_npix = npix;
        return gram_mfma_launch(Q, P.d_pack, wts, first, slots, s);
    }
    GramArgs g{};
    g.slc = slc;
    g.wts = wts;
    g.rows = P.rows;
    g.cols = P.cols;
    g.n = P.n;
    g.Nx = P.Nx;
    g.Ny = P.Ny;
    g.nul = P.nul;
    g.min_nb = P.min_nb;
    g.npx = P.npx;
    g.first = first;
    g.slots = slots;
    g.gamma = gamma;
    g.npix = npix;
    if (P.n <= 32) launch_gram_f64<32, 1, 32>(g, s);
    else if (P.n <= 64) launch_gram_f64<64, 1, 32>(g, s);
    else launch_gram_f64<64, 2, 16>(g, s);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int link_launch(const LinkPlan& P, const float2* slc, const uint32_t* wts, float2* phase,
                float* tcorr, float2* comp, hipStream_t s) {
    const long long first_out = (long long)P.row_begin * P.cols;
    const long long total = (long long)(P.row_end - P.row_begin) * P.cols;
    const bool fp64 = (P.precision == FRINGE_PREC_FP64);
    const bool mfma = link_uses_mfma(P.n, P.precision);
    if (mfma) {
        StageScope sc(STAGE_PACK, s);
        if (pack_launch(slc, P.rows, P.cols, P.n, P.precision != FRINGE_PREC_MIXED_BF16, P.d_pack, s) != 0)
            return -1;
    }
    for (long long done = 0; done < total; done += P.chunk) {
        const int slots = (int)((total - done) < P.chunk ? (total - done) : P.chunk);
        const long long first = first_out + done;
        {
        StageScope sc(STAGE_GRAM, s);
        if (!mfma) {
            GramArgs g{};
            g.slc = slc;
            g.wts = wts;
            g.rows = P.rows;
            g.cols = P.cols;
            g.n = P.n;
            g.Nx = P.Nx;
            g.Ny = P.Ny;
            g.nul = P.nul;
            g.min_nb = P.min_nb;
            g.npx = P.npx;
            g.first = first;
            g.slots = slots;
            g.gamma = P.d_gamma;
            g.npix = P.d_npix;
            if (P.n <= 32) launch_gram_f64<32, 1, 32>(g, s);
            else if (P.n <= 64) launch_gram_f64<64, 1, 32>(g, s);
            else launch_gram_f64<64, 2, 16>(g, s);
        } else {
            if (gram_mfma_launch(P, P.d_pack, wts, first, slots, s) != 0) return -1;
        }
        }
        SolveArgs a{};
        a.gamma = P.d_gamma;
        a.npix = P.d_npix;
        a.slc = slc;
        a.npx = P.npx;
        a.first = first;
        a.slots = slots;
        a.n = P.n;
        a.module = P.module;
        a.method = P.method;
        a.ref = P.ref;
        a.bw = P.bw;
        a.phase = phase;
        a.tcorr = tcorr;
        a.comp = comp;
        {
            StageScope sc(STAGE_SOLVE, s);
            if (fp64) dispatch_solve<double, double>(a, s);
            else if (mfma) dispatch_solve<float, float>(a, s);
            else dispatch_solve<float, double>(a, s);
        }
        if (hipGetLastError() != hipSuccess) return -1;
    }
    return 0;
}

}  // namespace fringe
