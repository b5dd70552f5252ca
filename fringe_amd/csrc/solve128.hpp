// solve128.hpp -- register-resident estimator chain for 64 < n <= 128 dates
// in FP32 (the mixed precisions), one pixel per four-wave workgroup.
//
// Thread t (0..255) holds half of row r = t / 2 of its pixel's matrices:
// the columns c = 2k + h (k < CH, h = t % 2), interleaved so that the live
// columns of every Householder step are split evenly between the two lanes
// of a row.  Values that belong to a row (the reflector entry, the product
// p = H v, the eigenvector entry) are held by both lanes of the row,
// bitwise alike: a lane pair combines its halves with one DPP swap, and
// a + b == b + a.  Sums over the rows reduce the even lanes of each wave by
// DPP and the four waves' partials through alternating LDS slots (one
// barrier each).  The FP64 inverse of |Gamma| keeps each lane's half row of
// |Gamma| in CH FP64 registers and applies the symmetric sweep operator one
// published pivot row at a time.  CH = 40 and 48 cover 80 and 96 dates at
// three waves per SIMD, CH = 56 and 64 cover 112 and 128 dates at two.
#pragma once
#include "solve32.hpp"

namespace fringe {

template <int CH> struct LRows {
    static constexpr int C = 2 * CH;   // columns (dates) held
    cx<float> R[C];     // published row
    cx<float> V[C];
    cx<float> W[C];
    cx<float> TAU[C];
    cx<float> OUT[C];
    cx<float> YREF;
    float D[C], E[C], E2[C];
    float B[C];
    float FP[C], FQ[C], FS[C], FX[C], FF[C], FK[C];
    double TR[C];       // sweep: published pivot row
    double PV;
    double red[2][4][2];   // [slot][wave][value]: block reduction partials
};

struct BRed4 {
    int seq = 0;
};

// ---- lane pairs and reductions over the four waves ------------------------
__device__ __forceinline__ float pswap(float x) { return __uint_as_float(dpp32<0xB1>(__float_as_uint(x))); }
__device__ __forceinline__ cx<float> pswap(cx<float> x) { return cx<float>{pswap(x.re), pswap(x.im)}; }
__device__ __forceinline__ cx<float> psum(cx<float> x) { return cadd(x, pswap(x)); }

template <typename Op, typename T, int CH> __device__ __forceinline__ T bred4(LRows<CH>& L, BRed4& br, T x) {
    x = redR<Op, 64>(x);
    const int s = br.seq++ & 1, w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) L.red[s][w][0] = (double)x;
    __syncthreads();
    return Op::f(Op::f((T)L.red[s][0][0], (T)L.red[s][1][0]), Op::f((T)L.red[s][2][0], (T)L.red[s][3][0]));
}
template <typename T, int CH> __device__ __forceinline__ cx<T> bsum4(LRows<CH>& L, BRed4& br, cx<T> x) {
    x.re = redR<OpAdd, 64>(x.re);
    x.im = redR<OpAdd, 64>(x.im);
    const int s = br.seq++ & 1, w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        L.red[s][w][0] = (double)x.re;
        L.red[s][w][1] = (double)x.im;
    }
    __syncthreads();
    return cx<T>{(T)((L.red[s][0][0] + L.red[s][1][0]) + (L.red[s][2][0] + L.red[s][3][0])),
                 (T)((L.red[s][0][1] + L.red[s][1][1]) + (L.red[s][2][1] + L.red[s][3][1]))};
}
// Index of the first thread (0..255) whose predicate holds, or 256.
template <int CH> __device__ __forceinline__ int bfirst4(LRows<CH>& L, BRed4& br, bool p) {
    const uint64_t b = __ballot(p);
    const int s = br.seq++ & 1, w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) L.red[s][w][0] = __builtin_bit_cast(double, b);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint64_t bq = __builtin_bit_cast(uint64_t, L.red[s][q][0]);
        if (bq) return 64 * q + (int)__builtin_ctzll(bq);
    }
    return 256;
}

// ---- FP64 inverse of |Gamma| - shift*I, half a row per thread -------------
// m[k] = |Gamma_{r,2k+h}| (off the diagonal), 1 - shift on the diagonal,
// identity outside n.
template <int CH>
__device__ __forceinline__ bool abs_gamma_half(double (&m)[CH], const cx<float>* grow, int n, int r, int hf, bool act,
                                               double shift) {
    bool fin = true;
#pragma unroll
    for (int k0 = 0; k0 < CH; k0 += 16) {
#pragma unroll
        for (int k = k0; k < (k0 + 16 < CH ? k0 + 16 : CH); ++k) {
            const int c = 2 * k + hf;
            double v = (c == r) ? 1.0 - shift : 0.0;
            if (act && r < n && c < n && c != r) {
                const cx<float> g = grow[c];
                v = (double)__builtin_amdgcn_sqrtf(g.re * g.re + g.im * g.im);   // one v_sqrt_f32 (~1 ulp)
                fin = fin && isfinite(v);
            }
            m[k] = v;
        }
        asm volatile("" ::: "memory");   // sixteen loads in flight at a time
    }
    return fin;
}

// Symmetric sweep operator over pivots 0..n-1 with the published pivot row
// P (P_k = A_kk - 1): every entry becomes A_rc - (P_r / p) P_c, right
// everywhere but at (k, k), which is left at 2 - 1/p; later pivots never
// read it, so subtracting 2 from each swept diagonal at the end gives
// -A^-1.  Returns whether every pivot was positive; pmin receives the
// smallest |pivot| up to the first one that is not.
template <int CH>
__device__ __forceinline__ bool sweep_half(LRows<CH>& L, double (&m)[CH], int n, int r, int hf, double& pmin) {
    constexpr int C = 2 * CH;
    bool ok = true;
    pmin = 1.0;
    for (int k = 0; k < n; ++k) {
        if (r == k) {
#pragma unroll
            for (int j = 0; j < CH; ++j) L.TR[2 * j + hf] = m[j];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const double p = L.TR[k];
            L.PV = p;
            L.TR[k] = p - 1.0;
        }
        __syncthreads();
        const double p = L.PV;
        if (ok) pmin = fmin(pmin, fabs(p));
        ok = ok && (p > 0.0);
        const double piv = fast_rcp(p);
        const double g = (r < n) ? L.TR[r < C ? r : C - 1] * piv : 0.0;
#pragma unroll
        for (int j0 = 0; j0 < CH; j0 += 16) {
#pragma unroll
            for (int j = j0; j < (j0 + 16 < CH ? j0 + 16 : CH); ++j) m[j] = fma(-g, L.TR[2 * j + hf], m[j]);
            asm volatile("" ::: "memory");
        }
        __syncthreads();   // every thread has read TR before the next pivot row replaces it
    }
#pragma unroll
    for (int j = 0; j < CH; ++j)
        if (2 * j + hf == r && r < n) m[j] -= 2.0;
    return ok;
}

// ---- Householder tridiagonalisation of the rows in h ----------------------
// Step j, for 2 K0 - 1 <= j <= 2 K0 + 6: local columns k < K0 (global
// columns below 2 K0) hold only finished entries and parked reflectors.
template <int CH, int K0>
__device__ __forceinline__ void tri_step_half(LRows<CH>& L, BRed4& br, cx<float> (&h)[CH], int j, int n, int r,
                                              int hf) {
    constexpr int C = 2 * CH;
    const int rc = r < C ? r : C - 1;
    const cx<float> alpha = cconj(L.R[j + 1]);
    const cx<float> xr = (r > j + 1 && r < n) ? cconj(L.R[rc]) : cx<float>{0.0f, 0.0f};
    const float xn2 = bred4<OpAdd>(L, br, hf ? 0.0f : cabs2(xr));
    const float ar = alpha.re, ai = alpha.im;
    cx<float> tau = {0.0f, 0.0f}, scal = {0.0f, 0.0f};
    float beta = ar;
    if (xn2 != 0.0f || ai != 0.0f) {
        const float nrm = __builtin_amdgcn_sqrtf(ar * ar + ai * ai + xn2);
        beta = (ar >= 0.0f) ? -nrm : nrm;
        const float ib = fast_rcp(beta);
        tau = cx<float>{(beta - ar) * ib, -ai * ib};
        const float dr_ = ar - beta;
        const float id = fast_rcp(dr_ * dr_ + ai * ai);
        scal = cx<float>{dr_ * id, -ai * id};
    }
    const cx<float> vr = (r == j + 1) ? cx<float>{1.0f, 0.0f}
                         : (r > j + 1 && r < n) ? cmul(xr, scal) : cx<float>{0.0f, 0.0f};
    if (hf == 0 && r < C) L.V[r] = vr;
    if (threadIdx.x == 0) {
        L.D[j] = L.R[j].re;
        L.E[j] = beta;
        L.TAU[j] = tau;
    }
    __syncthreads();
    cx<float> p = {0.0f, 0.0f};
#pragma unroll
    for (int k0 = (K0 / 8) * 8; k0 < CH; k0 += 8) {
#pragma unroll
        for (int k = (k0 > K0 ? k0 : K0); k < (k0 + 8 < CH ? k0 + 8 : CH); ++k) cfma(p, h[k], L.V[2 * k + hf]);
        asm volatile("" ::: "memory");
    }
    p = cmul(tau, psum(p));
    const cx<float> s = bsum4(L, br, hf ? cx<float>{0.0f, 0.0f} : cmulc(vr, p));
    const cx<float> a2 = cscale(cmul(tau, s), -0.5f);
    const cx<float> wr = (r > j && r < n) ? cadd(p, cmul(a2, vr)) : cx<float>{0.0f, 0.0f};
    if (hf == 0 && r < C) L.W[r] = wr;
    __syncthreads();
#pragma unroll
    for (int k0 = (K0 / 8) * 8; k0 < CH; k0 += 8) {
#pragma unroll
        for (int k = (k0 > K0 ? k0 : K0); k < (k0 + 8 < CH ? k0 + 8 : CH); ++k) {
            cfms_c(h[k], vr, L.W[2 * k + hf]);
            cfms_c(h[k], wr, L.V[2 * k + hf]);
        }
        asm volatile("" ::: "memory");
    }
    // park this row's entry of v_j in the column just annihilated (held by
    // the lane of parity j % 2, at k = j / 2)
#pragma unroll
    for (int k = (K0 > 0 ? K0 - 1 : 0); k < (K0 + 4 < CH ? K0 + 4 : CH); ++k)
        if (2 * k + hf == j && r > j + 1) h[k] = vr;
    if (r == j + 1) {
#pragma unroll
        for (int k = 0; k < CH; ++k) L.R[2 * k + hf] = h[k];
    }
    __syncthreads();
}

template <int CH, int K0>
__device__ __forceinline__ void tri_loop(LRows<CH>& L, BRed4& br, cx<float> (&h)[CH], int& j, int n, int r, int hf) {
    const int jend = (2 * K0 + 7 < n - 1) ? 2 * K0 + 7 : n - 1;
    for (; j < jend; ++j) tri_step_half<CH, K0>(L, br, h, j, n, r, hf);
    if constexpr (K0 + 4 < CH) tri_loop<CH, K0 + 4>(L, br, h, j, n, r, hf);
}

template <int CH>
__device__ __forceinline__ void tridiag_half(LRows<CH>& L, BRed4& br, cx<float> (&h)[CH], int n, int r, int hf) {
    if (r == 0) {
#pragma unroll
        for (int k = 0; k < CH; ++k) L.R[2 * k + hf] = h[k];
    }
    __syncthreads();
    int j = 0;
    tri_loop<CH, 0>(L, br, h, j, n, r, hf);
    if (threadIdx.x == 0) L.D[n - 1] = L.R[n - 1].re;
    __syncthreads();
}

// Sturm count of tridiag(L.D, L.E2) below sig.
template <int CH> __device__ __forceinline__ int sturm_below(const LRows<CH>& L, int n, float sig, float pivmin) {
    float q = L.D[0] - sig;
    if (fabsf(q) < pivmin) q = -pivmin;
    int cnt = q < 0.0f;
    for (int i = 1; i < n; ++i) {
        q = (L.D[i] - sig) - L.E2[i - 1] * fast_rcp(q);
        if (fabsf(q) < pivmin) q = -pivmin;
        cnt += q < 0.0f;
    }
    return cnt;
}

// ---- eigenvalue k (0-based ascending) of tridiag(D, E): 257-section -------
template <int CH>
__device__ __forceinline__ float bisect_half(LRows<CH>& L, BRed4& br, int n, int k, int r, int hf, float& tnorm,
                                             float& pm) {
    float lo = 3.0e38f, hi = -3.0e38f, em2 = 0.0f, tn = 0.0f;
    if (hf == 0 && r < n) {
        const float a = (r > 0) ? fabsf(L.E[r - 1]) : 0.0f;
        const float b = (r < n - 1) ? fabsf(L.E[r]) : 0.0f;
        lo = L.D[r] - a - b;
        hi = L.D[r] + a + b;
        tn = fabsf(L.D[r]) + a + b;
        if (r < n - 1) {
            L.E2[r] = L.E[r] * L.E[r];
            em2 = L.E2[r];
        }
    }
    lo = bred4<OpMin>(L, br, lo);
    hi = bred4<OpMax>(L, br, hi);
    em2 = bred4<OpMax>(L, br, em2);
    tn = bred4<OpMax>(L, br, tn);
    tnorm = tn;
    const float pivmin = Lim<float>::tiny() * max(1.0f, em2);
    pm = pivmin;
    const float pad = 4.0f * Lim<float>::eps() * tn + 2.0f * pivmin;
    lo -= pad;
    hi += pad;
    // four rounds of 257-section narrow the Gershgorin interval 4.4e9-fold
    const int t = threadIdx.x;
    for (int it = 0; it < 4; ++it) {
        const float w = (hi - lo) * (1.0f / 257.0f);
        const float sig = lo + w * (float)(t + 1);
        const int cnt = sturm_below(L, n, sig, pivmin);
        const int f = bfirst4(L, br, cnt > k);
        const float nhi = (f < 256) ? lo + w * (float)(f + 1) : hi;
        const float nlo = (f > 0) ? lo + w * (float)f : lo;
        lo = nlo;
        hi = nhi;
    }
    return 0.5f * (lo + hi);
}

// ---- eigenvector of tridiag(D, E) for lam: two rounds of inverse iteration
template <int CH>
__device__ __forceinline__ float tri_eigvec_half(LRows<CH>& L, BRed4& br, int n, int r, int hf, float lam,
                                                 float tnorm) {
    constexpr int C = 2 * CH;
    const float pert = max(Lim<float>::eps() * tnorm, Lim<float>::tiny());
    if (hf == 0 && r < C) L.B[r] = 0.5f + (float)((r * 37 + 11) & 63) * (1.0f / 64.0f);
    __syncthreads();
    float xme = 0.0f;
    float cd = L.D[0] - lam;
    for (int round = 0; round < 2; ++round) {
        float cx_ = L.B[0];
        if (round == 0) {
            float cu = (n > 1) ? L.E[0] : 0.0f;
            for (int i = 0; i < n - 1; ++i) {
                const float e = L.E[i];
                const float dn = L.D[i + 1] - lam;
                const float un = (i + 1 < n - 1) ? L.E[i + 1] : 0.0f;
                const float bn = L.B[i + 1];
                const bool keep = fabsf(cd) >= fabsf(e);
                float pv, qv, sv, xv, ncd, ncu, ncx, f;
                if (keep) {
                    const float cdp = (fabsf(cd) < pert) ? (cd < 0.0f ? -pert : pert) : cd;
                    f = e * fast_rcp(cdp);
                    pv = cdp; qv = cu; sv = 0.0f; xv = cx_;
                    ncd = dn - f * cu; ncu = un; ncx = bn - f * cx_;
                } else {
                    f = cd * fast_rcp(e);
                    pv = e; qv = dn; sv = un; xv = bn;
                    ncd = cu - f * dn; ncu = -f * un; ncx = cx_ - f * bn;
                }
                if (threadIdx.x == 0) {
                    L.FP[i] = pv;
                    L.FQ[i] = qv;
                    L.FS[i] = sv;
                    L.FX[i] = xv;
                    L.FF[i] = f;
                    L.FK[i] = keep ? 1.0f : 0.0f;
                }
                cd = ncd;
                cu = ncu;
                cx_ = ncx;
            }
            if (fabsf(cd) < pert) cd = (cd < 0.0f) ? -pert : pert;
        } else {
            for (int i = 0; i < n - 1; ++i) {
                const float bn = L.B[i + 1];
                const float f = L.FF[i];
                const bool keep = L.FK[i] != 0.0f;
                const float xv = keep ? cx_ : bn;
                const float ncx = keep ? bn - f * cx_ : cx_ - f * bn;
                if (threadIdx.x == 0) L.FX[i] = xv;
                cx_ = ncx;
            }
        }
        __syncthreads();
        float x1 = cx_ * fast_rcp(cd);
        float x2 = 0.0f;
        if (r == n - 1) xme = x1;
        for (int i = n - 2; i >= 0; --i) {
            const float xi = (L.FX[i] - L.FQ[i] * x1 - L.FS[i] * x2) * fast_rcp(L.FP[i]);
            if (r == i) xme = xi;
            x2 = x1;
            x1 = xi;
        }
        const float mx = bred4<OpMax>(L, br, (r < n) ? fabsf(xme) : 0.0f);
        const float sc = (mx > 0.0f && isfinite(mx)) ? 1.0f / mx : 1.0f;
        xme *= sc;
        if (hf == 0 && r < C) L.B[r] = xme;
        __syncthreads();
    }
    return (r < n) ? xme : 0.0f;
}

// y <- Q y with the reflectors parked in h (entry r > j+1 of v_j at column j).
template <int CH>
__device__ __forceinline__ cx<float> back_transform_half(LRows<CH>& L, BRed4& br, const cx<float> (&h)[CH], int n,
                                                         int r, int hf, cx<float> y) {
    constexpr int C = 2 * CH;
    // j is a compile-time index into h in every step: two fully unrolled
    // loops (one of 127 steps exceeds the unroller's size limit, and h would
    // then be indexed at run time, in scratch memory)
    auto step = [&](int j) {
        const cx<float> part = (hf == (j & 1) && r > j + 1 && r < n) ? h[j >> 1] : cx<float>{0.0f, 0.0f};
        const cx<float> v = (r == j + 1) ? cx<float>{1.0f, 0.0f} : psum(part);
        const cx<float> s = bsum4(L, br, hf ? cx<float>{0.0f, 0.0f} : cmulc(y, v));
        y = csub(y, cmul(v, cmul(L.TAU[j], s)));
    };
#pragma unroll
    for (int j = C - 2; j >= C / 2; --j)
        if (j <= n - 2) step(j);   // uniform
#pragma unroll
    for (int j = C / 2 - 1; j >= 0; --j)
        if (j <= n - 2) step(j);
    return y;
}

// Certification of the eigenpair (criteria at kCertScale).
template <int CH>
__device__ __forceinline__ bool eig_uncertain_half(LRows<CH>& L, BRed4& br, int n, int r, int hf, float lam, float tn,
                                                   float pivmin, bool largest, cx<float> y) {
    const float a2 = (r < n) ? cabs2(y) : 0.0f;
    const float s2 = bred4<OpAdd>(L, br, hf ? 0.0f : a2);
    const float m2 = bred4<OpMin>(L, br, (r < n) ? a2 : 3.0e38f);
    const float vmin = sqrtf(m2 / s2);
    const float delta = kCertScale * Lim<float>::eps() * tn / vmin;
    if (!(delta <= tn)) return true;
    const int c = sturm_below(L, n, largest ? lam - delta : lam + delta, pivmin);
    return largest ? (c <= n - 2) : (c >= 2);
}

template <int CH>
__device__ __forceinline__ bool load_half(LRows<CH>& L, BRed4& br, cx<float> (&h)[CH], const cx<float>* grow, int n,
                                          int r, int hf, bool act, int band) {
    bool fin = true;
#pragma unroll
    for (int k0 = 0; k0 < CH; k0 += 16) {
#pragma unroll
        for (int k = k0; k < (k0 + 16 < CH ? k0 + 16 : CH); ++k) {
            const int c = 2 * k + hf;
            cx<float> v = {0.0f, 0.0f};
            if (r < n && c < n) {
                if (act) v = grow[c];
                else if (c == r) v = cx<float>{1.0f, 0.0f};
                if (band >= 0 && (c - r > band || r - c > band)) v = cx<float>{0.0f, 0.0f};
            }
            fin = fin && isfinite(v.re) && isfinite(v.im);
            h[k] = v;
        }
        asm volatile("" ::: "memory");
    }
    return bred4<OpMin>(L, br, fin ? 1.0f : 0.0f) > 0.5f;
}

template <int CH>
__device__ __forceinline__ bool hadamard_half(LRows<CH>& L, BRed4& br, cx<float> (&h)[CH], const double (&m)[CH],
                                              const cx<float>* grow, int n, int r, int hf, bool act) {
    bool fin = true;
#pragma unroll
    for (int k0 = 0; k0 < CH; k0 += 16) {
#pragma unroll
        for (int k = k0; k < (k0 + 16 < CH ? k0 + 16 : CH); ++k) {
            const int c = 2 * k + hf;
            cx<float> v = {0.0f, 0.0f};
            if (r < n && c < n) {
                cx<float> g = (c == r) ? cx<float>{1.0f, 0.0f} : cx<float>{0.0f, 0.0f};
                if (act && c != r) g = grow[c];
                const float mm = (float)(-m[k]);   // m holds -|Gamma|^-1
                v = cx<float>{mm * g.re, mm * g.im};
                fin = fin && isfinite(v.re + v.im);
            }
            h[k] = v;
        }
        asm volatile("" ::: "memory");
    }
    return bred4<OpMin>(L, br, fin ? 1.0f : 0.0f) > 0.5f;
}

// One pixel per 256-thread workgroup; slot s of the round is pixel first + s.
// The FP32 result is certified; uncertified pixels go to the FP64 re-solve.
template <int CH, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CH <= 48 ? 3 : 2)))
void solve_rows_kernel(SolveArgs a) {
    constexpr int C = 2 * CH;
    __shared__ LRows<CH> L;
    BRed4 br;
    const int r = threadIdx.x >> 1, hf = threadIdx.x & 1;
    const int slot = (int)(gridDim.x - 1 - blockIdx.x);   // last-first, as the other solvers
    if (slot >= a.slots) return;                            // the whole workgroup
    const int n = a.n;
    const int np = a.npix[slot];
    const bool act = np >= 0;
    const cx<float>* gam = (const cx<float>*)a.gamma + (size_t)slot * n * n;   // the MFMA Gram's complex64
    const cx<float>* grow = gam + (size_t)min(r, n - 1) * n;
    const long long pix = a.first + slot;

    float code = 0.0f;
    bool failed = false;
    cx<float> h[CH];
    cx<float> y = {0.0f, 0.0f};
    float tn, pm;
    bool unc = false;

    if constexpr (MODE == MODE_MLE) {
        const bool fin = load_half(L, br, h, grow, n, r, hf, act, -1);
        tridiag_half(L, br, h, n, r, hf);
        const float lmin = bisect_half(L, br, n, 0, r, hf, tn, pm);
        if (!fin) { code = -1.0f; failed = true; }
        else if (!(lmin >= 1.0e-6f)) { code = -2.0f; failed = true; }
        unc = !(lmin >= 1.0e-6f + kCertThrMargin * Lim<float>::eps() * tn);
        double m[CH];
        double pmin;
        abs_gamma_half(m, grow, n, r, hf, act, 1.0e-6);
        const bool pd_shift = sweep_half(L, m, n, r, hf, pmin);
        if (!failed && !pd_shift) { code = -4.0f; failed = true; }
        unc = unc || !(pmin >= kCertPivotMin);
        abs_gamma_half(m, grow, n, r, hf, act, 0.0);
        const bool pd = sweep_half(L, m, n, r, hf, pmin);
        if (!failed && !pd) { code = -5.0f; failed = true; }
        unc = unc || !(pmin >= kCertPivotMin);
        const bool hfin = hadamard_half(L, br, h, m, grow, n, r, hf, act);
        tridiag_half(L, br, h, n, r, hf);
        const float lam = bisect_half(L, br, n, 0, r, hf, tn, pm);
        const float yr = tri_eigvec_half(L, br, n, r, hf, lam, tn);
        y = back_transform_half(L, br, h, n, r, hf, cx<float>{yr, 0.0f});
        if (!failed && !hfin) { code = -6.0f; failed = true; }
        if (!failed && !(lam >= 1.0e-6f)) { code = -7.0f; failed = true; }
        unc = unc || !(lam >= 1.0e-6f + kCertThrMargin * Lim<float>::eps() * tn);
        const bool ue = eig_uncertain_half(L, br, n, r, hf, lam, tn, pm, false, y);
        unc = fin && (unc || (!failed && ue));
    } else if constexpr (MODE == MODE_EVD) {
        const int band = (a.method == FRINGE_METHOD_STBAS) ? a.bw : -1;
        const bool fin = load_half(L, br, h, grow, n, r, hf, act, band);
        tridiag_half(L, br, h, n, r, hf);
        const float lam = bisect_half(L, br, n, n - 1, r, hf, tn, pm);
        const float yr = tri_eigvec_half(L, br, n, r, hf, lam, tn);
        y = back_transform_half(L, br, h, n, r, hf, cx<float>{yr, 0.0f});
        if (!fin) { code = -6.0f; failed = true; }
        else if (!(lam >= 1.0e-6f)) { code = -7.0f; failed = true; }
        const bool ue = eig_uncertain_half(L, br, n, r, hf, lam, tn, pm, true, y);
        unc = fin && (ue || !(lam >= 1.0e-6f + kCertThrMargin * Lim<float>::eps() * tn));
    } else {
        double m[CH];
        double pmin;
        bool fin = abs_gamma_half(m, grow, n, r, hf, act, 0.0);
        fin = bred4<OpMin>(L, br, fin ? 1.0f : 0.0f) > 0.5f;
        const bool pd = sweep_half(L, m, n, r, hf, pmin);
        const bool hfin = hadamard_half(L, br, h, m, grow, n, r, hf, act);
        const bool run_evd = !pd || !hfin;   // uniform over the workgroup
        tridiag_half(L, br, h, n, r, hf);
        const float lam = bisect_half(L, br, n, 0, r, hf, tn, pm);
        const float yr = tri_eigvec_half(L, br, n, r, hf, lam, tn);
        y = back_transform_half(L, br, h, n, r, hf, cx<float>{yr, 0.0f});
        const bool ue = eig_uncertain_half(L, br, n, r, hf, lam, tn, pm, false, y);
        unc = !(pmin >= kCertPivotMin) || (!run_evd && ue);
        if (act && fin && run_evd) {   // uniform: the whole workgroup is one pixel
            load_half(L, br, h, grow, n, r, hf, act, -1);
            tridiag_half(L, br, h, n, r, hf);
            const float le = bisect_half(L, br, n, n - 1, r, hf, tn, pm);
            const float ye = tri_eigvec_half(L, br, n, r, hf, le, tn);
            y = back_transform_half(L, br, h, n, r, hf, cx<float>{ye, 0.0f});
            unc = unc || eig_uncertain_half(L, br, n, r, hf, le, tn, pm, true, y);
        }
        if (!fin) { code = -1.0f; failed = true; }
        unc = unc && fin;
    }

    // ---- referenced phases (evd.cpp:738-749) ----
    if (hf == 0 && r == a.ref) L.YREF = y;
    __syncthreads();
    const cx<float> vref = L.YREF;
    bool okr, oku, okv;
    const cx<float> cref = unit_or(cconj(vref), cx<float>{1.0f, 0.0f}, okr);
    const cx<float> u = unit_or(cmulc(y, vref), cx<float>{0.0f, 0.0f}, oku);
    cx<float> o = oku ? u : (okr ? cref : unit_or(y, cx<float>{1.0f, 0.0f}, okv));
    if (r == a.ref) o = cx<float>{1.0f, 0.0f};
    if (hf == 0 && r < C) L.OUT[r] = o;
    __syncthreads();

    const bool write_ok = act && !failed;
    // ---- compressed SLC (evd.cpp:755-762) ----
    cx<double> cs = {0.0, 0.0};
    if (hf == 0 && r >= a.ref && r < n && write_ok) {
        const float2 z = a.slc[(long long)r * a.npx + pix];
        cs.re = (double)(z.x * o.re + z.y * o.im);
        cs.im = (double)(z.y * o.re - z.x * o.im);
    }
    cs.re = bred4<OpAdd>(L, br, cs.re);
    cs.im = bred4<OpAdd>(L, br, cs.im);
    // ---- temporal coherence (evd.cpp:770-786), pairs dealt out circulantly ----
    const bool stbas = (a.method == FRINGE_METHOD_STBAS);
    cx<float> tc = {0.0f, 0.0f};
    if (hf == 0 && r < n && write_ok) {
        bool okx;
        const cx<float> ci = unit_or(cx<float>{o.re, -o.im}, cx<float>{1.0f, 0.0f}, okx);
        const int half = (n - 1) / 2;
        const int dmax = ((n & 1) == 0 && r < n / 2) ? half + 1 : half;
        int jd = r;
        for (int d = 1; d <= dmax; ++d) {
            jd = jd + 1 == n ? 0 : jd + 1;
            const int lo = min(r, jd), hi = max(r, jd);
            if (stbas && hi - lo > a.bw) continue;
            const cx<float> gv = grow[jd];
            const cx<float> ph = unit_or(cx<float>{gv.re, gv.im}, cx<float>{1.0f, 0.0f}, okx);
            const cx<float> oj = L.OUT[jd];
            cx<float> t = cmul(cmul(ph, ci), oj);
            if (r > jd) t = cconj(t);
            tc = cadd(tc, t);
        }
    }
    tc = bsum4(L, br, tc);

    if (hf == 0 && r < n)
        a.phase[(long long)r * a.npx + pix] = write_ok ? make_float2(o.re, o.im) : make_float2(0.0f, 0.0f);
    if (threadIdx.x == 0) {
        long long counter = 0;
        if (stbas) {
            for (int i = 0; i < n; ++i) counter += min(n, i + a.bw + 1) - i - 1;
        } else {
            counter = (long long)n * (n - 1) / 2;
        }
        float tco = 0.0f;
        float2 co = make_float2(0.0f, 0.0f);
        if (act && failed) {
            tco = code;
        } else if (write_ok) {
            const double tre = (double)tc.re, tim = (double)tc.im;
            tco = (float)(sqrt(tre * tre + tim * tim) / (double)counter);
            const double inv = 1.0 / (double)(n - a.ref);
            co = make_float2((float)(cs.re * inv), (float)(cs.im * inv));
        }
        a.tcorr[pix] = tco;
        a.comp[pix] = co;
        if (act && unc) {
            const int e = atomicAdd(a.cert_count, 1);
            if (e < a.cert_cap) a.cert_list[e] = (int)pix;
        }
    }
}

}  // namespace fringe
