// solve96.hpp -- register-resident estimator chain for 64 < n <= 96 dates
// in FP32 (the mixed precisions), one pixel per two-wave workgroup.
//
// Lane r (0..127) owns row r of every matrix of its pixel in registers
// (96 columns; rows >= n idle), as the <= 64-date solver does within one
// wave.  Reductions over the pixel's rows combine the two waves' DPP results
// through a small LDS exchange (one barrier each); rows, reflectors and
// pivot rows are published through LDS.  The FP64 inverse of |Gamma| keeps
// each lane's row of |Gamma| in 96 FP64 registers and applies the symmetric
// sweep operator one published pivot row at a time, so nothing of the
// matrix lives in LDS.  A lane's row takes the whole register file (256
// VGPRs plus accumulation registers), one wave per SIMD: two pixels per
// compute unit, where the LDS-resident solver fits one at 96 dates, and no
// LDS round trip per matrix entry.
#pragma once
#include "solve32.hpp"

namespace fringe {

constexpr int kC96 = 96;   // columns per lane register row

struct L96 {
    cx<float> R[kC96];    // published row
    cx<float> V[kC96];
    cx<float> W[kC96];
    cx<float> TAU[kC96];
    cx<float> YREF;
    float D[kC96], E[kC96], E2[kC96];
    float B[kC96];
    float FP[kC96], FQ[kC96], FS[kC96], FX[kC96], FF[kC96], FK[kC96];
    double TR[kC96];      // sweep: published pivot row
    double PV;
    cx<float> OUT[kC96];
    double red[2][2][2];  // [slot][wave][value]: block reduction partials
    uint64_t bal[2];
};

// ---- reductions over the two waves of the workgroup -----------------------
// Each wave reduces its 64 lanes by DPP and permlane swaps; the two partials
// meet in LDS.  Partial slots alternate, so one barrier per reduction keeps a
// wave from overwriting a slot the other is still reading.
struct BRed {
    int seq = 0;
};
template <typename Op> __device__ __forceinline__ float wave64_red(float x) { return redR<Op, 64>(x); }
template <typename Op> __device__ __forceinline__ double wave64_red(double x) { return redR<Op, 64>(x); }

template <typename Op, typename T> __device__ __forceinline__ T bred(L96& L, BRed& br, T x) {
    x = wave64_red<Op>(x);
    const int s = br.seq++ & 1, w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) L.red[s][w][0] = (double)x;
    __syncthreads();
    return Op::f((T)L.red[s][0][0], (T)L.red[s][1][0]);
}
template <typename T> __device__ __forceinline__ cx<T> bsum(L96& L, BRed& br, cx<T> x) {
    x.re = wave64_red<OpAdd>(x.re);
    x.im = wave64_red<OpAdd>(x.im);
    const int s = br.seq++ & 1, w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        L.red[s][w][0] = (double)x.re;
        L.red[s][w][1] = (double)x.im;
    }
    __syncthreads();
    return cx<T>{(T)(L.red[s][0][0] + L.red[s][1][0]), (T)(L.red[s][0][1] + L.red[s][1][1])};
}
// Index of the first lane (0..127) whose predicate holds, or 128.
__device__ __forceinline__ int bfirst(L96& L, BRed& br, bool p) {
    const uint64_t b = __ballot(p);
    const int s = br.seq++ & 1, w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) L.red[s][w][0] = __builtin_bit_cast(double, b);
    __syncthreads();
    const uint64_t b0 = __builtin_bit_cast(uint64_t, L.red[s][0][0]);
    const uint64_t b1 = __builtin_bit_cast(uint64_t, L.red[s][1][0]);
    return b0 ? (int)__builtin_ctzll(b0) : (b1 ? 64 + (int)__builtin_ctzll(b1) : 128);
}

// ---- FP64 inverse of |Gamma| - shift*I, one row per lane ------------------
// m[c] = |Gamma_rc| (c != r), 1 - shift on the diagonal, identity outside n.
__device__ __forceinline__ bool abs_gamma_row96(double (&m)[kC96], const cx<float>* grow, int n, int r, bool act,
                                                double shift) {
    bool fin = true;
#pragma unroll
    for (int c0 = 0; c0 < kC96; c0 += 16) {
#pragma unroll
        for (int c = c0; c < c0 + 16; ++c) {
            double v = (c == r) ? 1.0 - shift : 0.0;
            if (act && r < n && c < n && c != r) {
                const cx<float> g = grow[c];
                v = (double)__builtin_amdgcn_sqrtf(g.re * g.re + g.im * g.im);   // one v_sqrt_f32 (~1 ulp)
                fin = fin && isfinite(v);
            }
            m[c] = v;
        }
        asm volatile("" ::: "memory");   // sixteen loads in flight at a time
    }
    return fin;
}

// Symmetric sweep operator over pivots 0..n-1 (see sweep_block): with the
// published pivot row P (P_k = A_kk - 1) every entry becomes
// A_rc - (P_r / p) P_c, which is right everywhere but at (k, k), where it
// leaves 2 - 1/p instead of -1/p.  Later pivots only add terms built from
// other entries to (k, k) and never read it, so subtracting 2 from every
// swept diagonal at the end gives -A^-1 exactly as the per-pivot fix would.
// Returns whether every pivot was positive; pmin receives the smallest
// |pivot| up to the first one that is not.
__device__ __forceinline__ bool sweep_rows96(L96& L, double (&m)[kC96], int n, int r, double& pmin) {
    bool ok = true;
    pmin = 1.0;
    for (int k = 0; k < n; ++k) {
        if (r == k) {
#pragma unroll
            for (int c = 0; c < kC96; ++c) L.TR[c] = m[c];
        }
        __syncthreads();
        if (r == 0) {
            const double p = L.TR[k];
            L.PV = p;
            L.TR[k] = p - 1.0;
        }
        __syncthreads();
        const double p = L.PV;
        if (ok) pmin = fmin(pmin, fabs(p));
        ok = ok && (p > 0.0);
        const double piv = fast_rcp(p);
        const double g = (r < n) ? L.TR[r] * piv : 0.0;
#pragma unroll
        for (int c0 = 0; c0 < kC96; c0 += 16) {
#pragma unroll
            for (int c = c0; c < c0 + 16; ++c) m[c] = fma(-g, L.TR[c], m[c]);
            asm volatile("" ::: "memory");
        }
        __syncthreads();   // every lane has read TR before the next pivot row replaces it
    }
#pragma unroll
    for (int c = 0; c < kC96; ++c)
        if (c == r && r < n) m[c] -= 2.0;
    return ok;
}

// ---- Householder tridiagonalisation of the rows in h ----------------------
template <int C0, int STEP>
__device__ __forceinline__ void tri_step96(L96& L, BRed& br, cx<float> (&h)[kC96], int j, int n, int r) {
    const cx<float> alpha = cconj(L.R[j + 1]);
    const cx<float> xr = (r > j + 1 && r < n) ? cconj(L.R[r]) : cx<float>{0.0f, 0.0f};
    const float xn2 = bred<OpAdd>(L, br, cabs2(xr));
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
    if (r < kC96) L.V[r] = vr;
    if (r == 0) {
        L.D[j] = L.R[j].re;
        L.E[j] = beta;
        L.TAU[j] = tau;
    }
    __syncthreads();
    // column chunks of 16 with a compiler barrier between them: the LDS
    // operands of one chunk are loaded just before use, not all 96 at once
    cx<float> p = {0.0f, 0.0f};
#pragma unroll
    for (int c0 = (C0 / 16) * 16; c0 < kC96; c0 += 16) {
#pragma unroll
        for (int c = (c0 > C0 ? c0 : C0); c < c0 + 16; ++c) cfma(p, h[c], L.V[c]);
        asm volatile("" ::: "memory");
    }
    p = cmul(tau, p);
    const cx<float> s = bsum(L, br, cmulc(vr, p));
    const cx<float> a2 = cscale(cmul(tau, s), -0.5f);
    const cx<float> wr = (r > j && r < n) ? cadd(p, cmul(a2, vr)) : cx<float>{0.0f, 0.0f};
    if (r < kC96) L.W[r] = wr;
    __syncthreads();
#pragma unroll
    for (int c0 = (C0 / 16) * 16; c0 < kC96; c0 += 16) {
#pragma unroll
        for (int c = (c0 > C0 ? c0 : C0); c < c0 + 16; ++c) {
            cfms_c(h[c], vr, L.W[c]);
            cfms_c(h[c], wr, L.V[c]);
        }
        asm volatile("" ::: "memory");
    }
    // park this lane's entry of v_j in the column just annihilated
#pragma unroll
    for (int c = (C0 > 0 ? C0 - 1 : 0); c < (C0 + STEP - 1 < kC96 ? C0 + STEP - 1 : kC96); ++c)
        if (c == j && r > j + 1) h[c] = vr;
    if (r == j + 1) {
#pragma unroll
        for (int c = 0; c < kC96; ++c) L.R[c] = h[c];
    }
    __syncthreads();
}

__device__ __forceinline__ void tridiag96(L96& L, BRed& br, cx<float> (&h)[kC96], int n, int r) {
    if (r == 0) {
#pragma unroll
        for (int c = 0; c < kC96; ++c) L.R[c] = h[c];
    }
    __syncthreads();
    int j = 0;
    for (; j < min(7, n - 1); ++j) tri_step96<0, 8>(L, br, h, j, n, r);
    for (; j < min(15, n - 1); ++j) tri_step96<8, 8>(L, br, h, j, n, r);
    for (; j < min(23, n - 1); ++j) tri_step96<16, 8>(L, br, h, j, n, r);
    for (; j < min(31, n - 1); ++j) tri_step96<24, 8>(L, br, h, j, n, r);
    for (; j < min(39, n - 1); ++j) tri_step96<32, 8>(L, br, h, j, n, r);
    for (; j < min(47, n - 1); ++j) tri_step96<40, 8>(L, br, h, j, n, r);
    for (; j < min(55, n - 1); ++j) tri_step96<48, 8>(L, br, h, j, n, r);
    for (; j < min(63, n - 1); ++j) tri_step96<56, 8>(L, br, h, j, n, r);
    for (; j < min(71, n - 1); ++j) tri_step96<64, 8>(L, br, h, j, n, r);
    for (; j < min(79, n - 1); ++j) tri_step96<72, 8>(L, br, h, j, n, r);
    for (; j < min(87, n - 1); ++j) tri_step96<80, 8>(L, br, h, j, n, r);
    for (; j < n - 1; ++j) tri_step96<88, 8>(L, br, h, j, n, r);
    if (r == 0) L.D[n - 1] = L.R[n - 1].re;
    __syncthreads();
}

// Sturm count of tridiag(L.D, L.E2) below sig (every lane alike).
__device__ __forceinline__ int sturm_below(const L96& L, int n, float sig, float pivmin) {
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

// ---- eigenvalue k (0-based ascending) of tridiag(D, E): 129-section -------
__device__ __forceinline__ float bisect96(L96& L, BRed& br, int n, int k, int r, float& tnorm, float& pm) {
    float lo = 3.0e38f, hi = -3.0e38f, em2 = 0.0f, tn = 0.0f;
    if (r < n) {
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
    lo = bred<OpMin>(L, br, lo);
    hi = bred<OpMax>(L, br, hi);
    em2 = bred<OpMax>(L, br, em2);
    tn = bred<OpMax>(L, br, tn);
    tnorm = tn;
    const float pivmin = Lim<float>::tiny() * max(1.0f, em2);
    pm = pivmin;
    const float pad = 4.0f * Lim<float>::eps() * tn + 2.0f * pivmin;
    lo -= pad;
    hi += pad;
    // four rounds of 129-section narrow the Gershgorin interval 2.8e8-fold
    for (int it = 0; it < 4; ++it) {
        const float w = (hi - lo) * (1.0f / 129.0f);
        const float sig = lo + w * (float)(r + 1);
        const int cnt = sturm_below(L, n, sig, pivmin);
        const int f = bfirst(L, br, cnt > k);
        const float nhi = (f < 128) ? lo + w * (float)(f + 1) : hi;
        const float nlo = (f > 0) ? lo + w * (float)f : lo;
        lo = nlo;
        hi = nhi;
    }
    return 0.5f * (lo + hi);
}

// ---- eigenvector of tridiag(D, E) for lam: two rounds of inverse iteration
__device__ __forceinline__ float tri_eigvec96(L96& L, BRed& br, int n, int r, float lam, float tnorm) {
    const float pert = max(Lim<float>::eps() * tnorm, Lim<float>::tiny());
    if (r < kC96) L.B[r] = 0.5f + (float)((r * 37 + 11) & 63) * (1.0f / 64.0f);
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
                if (r == 0) {
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
                if (r == 0) L.FX[i] = xv;
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
        const float mx = bred<OpMax>(L, br, (r < n) ? fabsf(xme) : 0.0f);
        const float sc = (mx > 0.0f && isfinite(mx)) ? 1.0f / mx : 1.0f;
        xme *= sc;
        if (r < kC96) L.B[r] = xme;
        __syncthreads();
    }
    return (r < n) ? xme : 0.0f;
}

// y <- Q y with the reflectors parked in h (entry r > j+1 of v_j in h[j]).
__device__ __forceinline__ cx<float> back_transform96(L96& L, BRed& br, const cx<float> (&h)[kC96], int n, int r,
                                                      cx<float> y) {
#pragma unroll
    for (int j = kC96 - 2; j >= 0; --j) {
        if (j > n - 2) continue;   // uniform
        const cx<float> v = (r == j + 1) ? cx<float>{1.0f, 0.0f}
                            : (r > j + 1 && r < n) ? h[j] : cx<float>{0.0f, 0.0f};
        const cx<float> s = bsum(L, br, cmulc(y, v));
        y = csub(y, cmul(v, cmul(L.TAU[j], s)));
    }
    return y;
}

// Certification of the eigenpair (criteria at kCertScale).
__device__ __forceinline__ bool eig_uncertain96(L96& L, BRed& br, int n, int r, float lam, float tn, float pivmin,
                                                bool largest, cx<float> y) {
    const float a2 = (r < n) ? cabs2(y) : 0.0f;
    const float s2 = bred<OpAdd>(L, br, a2);
    const float m2 = bred<OpMin>(L, br, (r < n) ? a2 : 3.0e38f);
    const float vmin = sqrtf(m2 / s2);
    const float delta = kCertScale * Lim<float>::eps() * tn / vmin;
    if (!(delta <= tn)) return true;
    const int c = sturm_below(L, n, largest ? lam - delta : lam + delta, pivmin);
    return largest ? (c <= n - 2) : (c >= 2);
}

__device__ __forceinline__ bool load_rows96(L96& L, BRed& br, cx<float> (&h)[kC96], const cx<float>* grow, int n,
                                            int r, bool act, int band) {
    bool fin = true;
#pragma unroll
    for (int c0 = 0; c0 < kC96; c0 += 16) {
#pragma unroll
        for (int c = c0; c < c0 + 16; ++c) {
            cx<float> v = {0.0f, 0.0f};
            if (r < n && c < n) {
                if (act) v = grow[c];
                else if (c == r) v = cx<float>{1.0f, 0.0f};
                if (band >= 0 && (c - r > band || r - c > band)) v = cx<float>{0.0f, 0.0f};
            }
            fin = fin && isfinite(v.re) && isfinite(v.im);
            h[c] = v;
        }
        asm volatile("" ::: "memory");
    }
    return bred<OpMin>(L, br, fin ? 1.0f : 0.0f) > 0.5f;
}

__device__ __forceinline__ bool hadamard_rows96(L96& L, BRed& br, cx<float> (&h)[kC96], const double (&m)[kC96],
                                                const cx<float>* grow, int n, int r, bool act) {
    bool fin = true;
#pragma unroll
    for (int c0 = 0; c0 < kC96; c0 += 16) {
#pragma unroll
        for (int c = c0; c < c0 + 16; ++c) {
            cx<float> v = {0.0f, 0.0f};
            if (r < n && c < n) {
                cx<float> g = (c == r) ? cx<float>{1.0f, 0.0f} : cx<float>{0.0f, 0.0f};
                if (act && c != r) g = grow[c];
                const float mm = (float)(-m[c]);   // m holds -|Gamma|^-1
                v = cx<float>{mm * g.re, mm * g.im};
                fin = fin && isfinite(v.re + v.im);
            }
            h[c] = v;
        }
        asm volatile("" ::: "memory");
    }
    return bred<OpMin>(L, br, fin ? 1.0f : 0.0f) > 0.5f;
}

// One pixel per 128-thread workgroup; slot s of the round is pixel first + s.
// The FP32 result is certified; uncertified pixels go to the FP64 re-solve.
template <int MODE>
__global__ __launch_bounds__(128) void solve96_kernel(SolveArgs a) {
    __shared__ L96 L;
    BRed br;
    const int r = threadIdx.x;
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
    cx<float> h[kC96];
    cx<float> y = {0.0f, 0.0f};
    float tn, pm;
    bool unc = false;

    if constexpr (MODE == MODE_MLE) {
        const bool fin = load_rows96(L, br, h, grow, n, r, act, -1);
        tridiag96(L, br, h, n, r);
        const float lmin = bisect96(L, br, n, 0, r, tn, pm);
        if (!fin) { code = -1.0f; failed = true; }
        else if (!(lmin >= 1.0e-6f)) { code = -2.0f; failed = true; }
        unc = !(lmin >= 1.0e-6f + kCertThrMargin * Lim<float>::eps() * tn);
        double m[kC96];
        double pmin;
        abs_gamma_row96(m, grow, n, r, act, 1.0e-6);
        const bool pd_shift = sweep_rows96(L, m, n, r, pmin);
        if (!failed && !pd_shift) { code = -4.0f; failed = true; }
        unc = unc || !(pmin >= kCertPivotMin);
        abs_gamma_row96(m, grow, n, r, act, 0.0);
        const bool pd = sweep_rows96(L, m, n, r, pmin);
        if (!failed && !pd) { code = -5.0f; failed = true; }
        unc = unc || !(pmin >= kCertPivotMin);
        const bool hfin = hadamard_rows96(L, br, h, m, grow, n, r, act);
        tridiag96(L, br, h, n, r);
        const float lam = bisect96(L, br, n, 0, r, tn, pm);
        const float yr = tri_eigvec96(L, br, n, r, lam, tn);
        y = back_transform96(L, br, h, n, r, cx<float>{yr, 0.0f});
        if (!failed && !hfin) { code = -6.0f; failed = true; }
        if (!failed && !(lam >= 1.0e-6f)) { code = -7.0f; failed = true; }
        unc = unc || !(lam >= 1.0e-6f + kCertThrMargin * Lim<float>::eps() * tn);
        const bool ue = eig_uncertain96(L, br, n, r, lam, tn, pm, false, y);
        unc = fin && (unc || (!failed && ue));
    } else if constexpr (MODE == MODE_EVD) {
        const int band = (a.method == FRINGE_METHOD_STBAS) ? a.bw : -1;
        const bool fin = load_rows96(L, br, h, grow, n, r, act, band);
        tridiag96(L, br, h, n, r);
        const float lam = bisect96(L, br, n, n - 1, r, tn, pm);
        const float yr = tri_eigvec96(L, br, n, r, lam, tn);
        y = back_transform96(L, br, h, n, r, cx<float>{yr, 0.0f});
        if (!fin) { code = -6.0f; failed = true; }
        else if (!(lam >= 1.0e-6f)) { code = -7.0f; failed = true; }
        const bool ue = eig_uncertain96(L, br, n, r, lam, tn, pm, true, y);
        unc = fin && (ue || !(lam >= 1.0e-6f + kCertThrMargin * Lim<float>::eps() * tn));
    } else {
        double m[kC96];
        double pmin;
        bool fin = abs_gamma_row96(m, grow, n, r, act, 0.0);
        fin = bred<OpMin>(L, br, fin ? 1.0f : 0.0f) > 0.5f;
        const bool pd = sweep_rows96(L, m, n, r, pmin);
        const bool hfin = hadamard_rows96(L, br, h, m, grow, n, r, act);
        const bool run_evd = !pd || !hfin;   // uniform over the workgroup
        tridiag96(L, br, h, n, r);
        const float lam = bisect96(L, br, n, 0, r, tn, pm);
        const float yr = tri_eigvec96(L, br, n, r, lam, tn);
        y = back_transform96(L, br, h, n, r, cx<float>{yr, 0.0f});
        const bool ue = eig_uncertain96(L, br, n, r, lam, tn, pm, false, y);
        unc = !(pmin >= kCertPivotMin) || (!run_evd && ue);
        if (act && fin && run_evd) {   // uniform: the whole workgroup is one pixel
            load_rows96(L, br, h, grow, n, r, act, -1);
            tridiag96(L, br, h, n, r);
            const float le = bisect96(L, br, n, n - 1, r, tn, pm);
            const float ye = tri_eigvec96(L, br, n, r, le, tn);
            y = back_transform96(L, br, h, n, r, cx<float>{ye, 0.0f});
            unc = unc || eig_uncertain96(L, br, n, r, le, tn, pm, true, y);
        }
        if (!fin) { code = -1.0f; failed = true; }
        unc = unc && fin;
    }

    // ---- referenced phases (evd.cpp:738-749) ----
    if (r == a.ref) L.YREF = y;
    __syncthreads();
    const cx<float> vref = L.YREF;
    bool okr, oku, okv;
    const cx<float> cref = unit_or(cconj(vref), cx<float>{1.0f, 0.0f}, okr);
    const cx<float> u = unit_or(cmulc(y, vref), cx<float>{0.0f, 0.0f}, oku);
    cx<float> o = oku ? u : (okr ? cref : unit_or(y, cx<float>{1.0f, 0.0f}, okv));
    if (r == a.ref) o = cx<float>{1.0f, 0.0f};
    if (r < kC96) L.OUT[r] = o;
    __syncthreads();

    const bool write_ok = act && !failed;
    // ---- compressed SLC (evd.cpp:755-762) ----
    cx<double> cs = {0.0, 0.0};
    if (r >= a.ref && r < n && write_ok) {
        const float2 z = a.slc[(long long)r * a.npx + pix];
        cs.re = (double)(z.x * o.re + z.y * o.im);
        cs.im = (double)(z.y * o.re - z.x * o.im);
    }
    cs.re = bred<OpAdd>(L, br, cs.re);
    cs.im = bred<OpAdd>(L, br, cs.im);
    // ---- temporal coherence (evd.cpp:770-786), pairs dealt out circulantly ----
    const bool stbas = (a.method == FRINGE_METHOD_STBAS);
    cx<float> tc = {0.0f, 0.0f};
    if (r < n && write_ok) {
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
            const cx<float> ph = unit_or(cx<float>{(float)gv.re, (float)gv.im}, cx<float>{1.0f, 0.0f}, okx);
            const cx<float> oj = L.OUT[jd];
            cx<float> t = cmul(cmul(ph, ci), oj);
            if (r > jd) t = cconj(t);
            tc = cadd(tc, t);
        }
    }
    tc = bsum(L, br, tc);

    if (r < n) a.phase[(long long)r * a.npx + pix] = write_ok ? make_float2(o.re, o.im) : make_float2(0.0f, 0.0f);
    if (r == 0) {
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
