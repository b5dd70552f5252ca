// gram64.hpp -- FP64 masked-neighbourhood Gram of one pixel per lane group
// (the plain GPU path's coherence, evd.cpp:530-582 / phase_link.cpp:496-538).
//
// One group of G lanes per pixel; lane g owns rows g + G*q (q < R) and
// accumulates them against CB columns at a time.  The masked, in-image
// window positions are compacted into a list first, so the work is
// proportional to popcount(mask), not to the window.  Each product is the
// complex<float> sample product widened to double and accumulated in double,
// as the reference's loop does.
//
// Used by the chunked FP64 Gram kernel and by the FP64 re-solve of the
// mixed-precision path, which computes the coherence of a listed pixel into
// LDS and solves it in the same launch.
#pragma once
#include "device_common.hpp"

namespace fringe {

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

// LDS bytes of one group: the position list, one row of samples, the diagonal.
__host__ __device__ inline size_t gram_f64_group_bytes(int n, int Nx, int Ny, int CB) {
    const int Wn = (2 * Ny + 1) * (2 * Nx + 1);
    const int NC = (n + CB - 1) / CB;
    size_t b = ((size_t)Wn * 4 + 15) & ~(size_t)15;
    b += (size_t)NC * CB * sizeof(cx<double>);
    b += (size_t)n * sizeof(double);
    return (b + 15) & ~(size_t)15;
}

// Barrier of the Gram's lane groups: the workgroup's, or (WSYNC, a single
// 64-lane group called by one wave while the others wait) the wave's.
template <bool WSYNC> __device__ __forceinline__ void gram_sync() {
    if constexpr (WSYNC) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// Coherence of pixel `pix` into gam[n][n] (both triangles, unit diagonal);
// returns the neighbour count, or -1 when the pixel is not linked (centre
// bit clear, or fewer than min_nb neighbours; gam is then not written).
// Every group of the workgroup calls it together (it holds barriers);
// `valid` false makes a group take part without a pixel.  WSYNC (G = 64):
// the calling wave synchronises only itself, and computes the column blocks
// cb_first, cb_first + cb_step, ... (several waves can share a pixel's
// blocks; each entry is the same sum in the same order whatever the split,
// and every wave accumulates the diagonal in its first block).
template <int G, int R, int CB, bool WSYNC = false>
__device__ __forceinline__ int gram_f64_pixel(const GramArgs& a, bool valid, long long pix, cx<double>* gam,
                                              char* base, int cb_first = 0, int cb_step = 1) {
    static_assert(!WSYNC || G == 64, "a wave-synchronised Gram is one 64-lane group");
    constexpr int GPW = 64 / G;
    const int g = threadIdx.x % G;
    const int n = a.n;
    const int Wn = (2 * a.Ny + 1) * (2 * a.Nx + 1);
    const int NC = (n + CB - 1) / CB;
    int* list = (int*)base;
    cx<double>* zb = (cx<double>*)(base + ((Wn * 4 + 15) & ~15));
    double* dg = (double*)(zb + NC * CB);

    const int r0 = (int)(pix / a.cols), c0 = (int)(pix % a.cols);
    const uint32_t* w = a.wts + pix * a.nul;
    const bool cen = valid && shp_get(w, shp_flat(0, 0, a.Nx, a.Ny));
    const int cnt = build_position_list<G>(w, cen, r0, c0, a.rows, a.cols, a.Nx, a.Ny, list);
    const bool go = cen && cnt >= a.min_nb;
    const int K = go ? cnt : 0;
    int Kw = K;
    if (GPW > 1) Kw = max(Kw, __shfl_xor(Kw, 32));
    for (int c = g; c < NC * CB; c += G) zb[c] = cx<double>{0.0, 0.0};
    gram_sync<WSYNC>();

    for (int cb = cb_first; cb < NC; cb += cb_step) {
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
            gram_sync<WSYNC>();
#pragma unroll
            for (int q = 0; q < R; ++q) {
#pragma unroll
                for (int c = 0; c < CB; ++c) cfma_c(acc[q][c], z[q], zb[cb * CB + c]);
                if (cb == cb_first) d[q] += cabs2(z[q]);
            }
            gram_sync<WSYNC>();
        }
        if (cb == cb_first) {
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int r = g + G * q;
                if (r < n) dg[r] = d[q];
            }
            gram_sync<WSYNC>();
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
        gram_sync<WSYNC>();
    }
    return go ? cnt : -1;
}

}  // namespace fringe
