// solve128.hpp -- register-resident estimator chain for 32 < n <= 128 dates,
// one pixel per workgroup of RW rows (48 .. 128, the fewest that hold n) of
// Q threads each.
//
// Thread t holds a 1/Q share of row r = t / Q of its pixel's matrices: the
// columns c = Q k + q (k < CH, q = t % Q), interleaved so that the live
// columns of every Householder step are split evenly between the Q threads
// of a row.  Values that belong to a row (the reflector entry, the product
// p = H v, the eigenvector entry) are held by all Q threads of the row,
// bitwise alike: the threads combine their shares by DPP swaps within a
// quad, and each stage adds two equal partials commutatively.  Sums over the
// rows reduce the q = 0 threads of each wave by DPP and the waves' partials
// through alternating LDS slots (one barrier each).  The FP64 inverse of
// |Gamma| keeps each thread's share of its row of |Gamma| in CH FP64
// registers and applies the symmetric sweep operator one published pivot row
// at a time.
//
//   T = float, Q = 4 up to 48 dates (three waves), else Q = 2 (two to four
//   waves): the mixed precisions, coherence from the MFMA Gram, FP32 results
//   certified (uncertified pixels are listed for the FP64 re-solve).
//   CH = 12, 32 .. 64.
//   T = double, Q = 4 (three to eight waves): the plain FP64 path, and the
//   FP64 re-solve of the listed pixels (each workgroup computes the listed
//   pixel's FP64 coherence into its global area first), so that re-solved
//   pixels equal the FP64 path bit for bit.  CH = 12 .. 32.
#pragma once
#include "gram64.hpp"
#include "solve32.hpp"

namespace fringe {

template <typename T, int Q, int CH, int RW> struct LRows {
    static constexpr int C = Q * CH;        // columns (dates) held
    static constexpr int NW = RW * Q / 64;  // waves (RW rows of Q threads)
    cx<T> R[C];         // published row
    cx<T> V[C];
    cx<T> W[C];
    cx<T> TAU[C];
    cx<float> OUT[C];
    cx<T> YREF;
    T D[C], E[C], E2[C];
    T B[C];
    T FP[C], FQ[C], FS[C], FX[C], FF[C], FK[C];
    double TR[C];       // sweep: published pivot row
    double PV;
    double red[2][NW][2];   // [slot][wave][value]: block reduction partials
};

struct BRedN {
    int seq = 0;
};

// ---- the Q threads of a row and reductions over the waves ------------------
template <int CTRL> __device__ __forceinline__ float dppf(float x) { return __uint_as_float(dpp32<CTRL>(__float_as_uint(x))); }
template <int CTRL> __device__ __forceinline__ double dppf(double x) { return dpp64<CTRL>(x); }
// sum of the row's Q shares, the same bits in every thread of the row
template <int Q, typename T> __device__ __forceinline__ T qsum(T x) {
    x = x + dppf<0xB1>(x);                       // quad_perm [1,0,3,2]
    if constexpr (Q == 4) x = x + dppf<0x4E>(x);  // quad_perm [2,3,0,1]
    return x;
}
template <int Q, typename T> __device__ __forceinline__ cx<T> qsum(cx<T> x) {
    return cx<T>{qsum<Q>(x.re), qsum<Q>(x.im)};
}

template <typename Op, typename T, typename S, int Q, int CH, int RW>
__device__ __forceinline__ S bredn(LRows<T, Q, CH, RW>& L, BRedN& br, S x) {
    constexpr int NW = LRows<T, Q, CH, RW>::NW;
    x = redR<Op, 64>(x);
    const int s = br.seq++ & 1, w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) L.red[s][w][0] = (double)x;
    __syncthreads();
    S a = (S)L.red[s][0][0];
#pragma unroll
    for (int q = 1; q < NW; ++q) a = Op::f(a, (S)L.red[s][q][0]);
    return a;
}
template <typename T, typename S, int Q, int CH, int RW>
__device__ __forceinline__ cx<S> bsumn(LRows<T, Q, CH, RW>& L, BRedN& br, cx<S> x) {
    constexpr int NW = LRows<T, Q, CH, RW>::NW;
    x.re = redR<OpAdd, 64>(x.re);
    x.im = redR<OpAdd, 64>(x.im);
    const int s = br.seq++ & 1, w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        L.red[s][w][0] = (double)x.re;
        L.red[s][w][1] = (double)x.im;
    }
    __syncthreads();
    double re = L.red[s][0][0], im = L.red[s][0][1];
#pragma unroll
    for (int q = 1; q < NW; ++q) {
        re += L.red[s][q][0];
        im += L.red[s][q][1];
    }
    return cx<S>{(S)re, (S)im};
}
// Index of the first thread whose predicate holds, or 64 NW.
template <typename T, int Q, int CH, int RW>
__device__ __forceinline__ int bfirstn(LRows<T, Q, CH, RW>& L, BRedN& br, bool p) {
    constexpr int NW = LRows<T, Q, CH, RW>::NW;
    const uint64_t b = __ballot(p);
    const int s = br.seq++ & 1, w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) L.red[s][w][0] = __builtin_bit_cast(double, b);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NW; ++q) {
        const uint64_t bq = __builtin_bit_cast(uint64_t, L.red[s][q][0]);
        if (bq) return 64 * q + (int)__builtin_ctzll(bq);
    }
    return 64 * NW;
}

template <typename T> __device__ __forceinline__ T sqrt_t(T x) {
    if constexpr (sizeof(T) == 4) return __builtin_amdgcn_sqrtf(x);   // one v_sqrt_f32 (~1 ulp)
    else return sqrt(x);
}

// ---- FP64 inverse of |Gamma| - shift*I, a share of a row per thread --------
// m[k] = |Gamma_{r,Qk+q}| (off the diagonal), 1 - shift on the diagonal,
// identity outside n.
template <int Q, int CH, typename Tg>
__device__ __forceinline__ bool abs_gamma_share(double (&m)[CH], const cx<Tg>* grow, int n, int r, int q, bool act,
                                                double shift) {
    bool fin = true;
#pragma unroll
    for (int k0 = 0; k0 < CH; k0 += 16) {
#pragma unroll
        for (int k = k0; k < (k0 + 16 < CH ? k0 + 16 : CH); ++k) {
            const int c = Q * k + q;
            double v = (c == r) ? 1.0 - shift : 0.0;
            if (act && r < n && c < n && c != r) {
                const cx<Tg> g = grow[c];
                v = (double)sqrt_t<Tg>(g.re * g.re + g.im * g.im);
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
template <typename T, int Q, int CH, int RW>
__device__ __forceinline__ bool sweep_share(LRows<T, Q, CH, RW>& L, double (&m)[CH], int n, int r, int q, double& pmin) {
    constexpr int C = Q * CH;
    bool ok = true;
    pmin = 1.0;
    for (int k = 0; k < n; ++k) {
        if (r == k) {
#pragma unroll
            for (int j = 0; j < CH; ++j) L.TR[Q * j + q] = m[j];
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
            for (int j = j0; j < (j0 + 16 < CH ? j0 + 16 : CH); ++j) m[j] = fma(-g, L.TR[Q * j + q], m[j]);
            asm volatile("" ::: "memory");
        }
        __syncthreads();   // every thread has read TR before the next pivot row replaces it
    }
#pragma unroll
    for (int j = 0; j < CH; ++j)
        if (Q * j + q == r && r < n) m[j] -= 2.0;
    return ok;
}

// ---- Householder tridiagonalisation of the rows in h ----------------------
// Step j, for Q K0 - 1 <= j <= Q K0 + 6: local columns k < K0 (global columns
// below Q K0, a multiple of 8) hold only finished entries and parked
// reflectors.
template <typename T, int Q, int CH, int RW, int K0>
__device__ __forceinline__ void tri_step_share(LRows<T, Q, CH, RW>& L, BRedN& br, cx<T> (&h)[CH], int j, int n, int r,
                                               int q) {
    constexpr int C = Q * CH;
    constexpr int KB = 8 / Q;   // local columns per 8 global ones
    const int rc = r < C ? r : C - 1;
    const cx<T> alpha = cconj(L.R[j + 1]);
    const cx<T> xr = (r > j + 1 && r < n) ? cconj(L.R[rc]) : cx<T>{T(0), T(0)};
    const T xn2 = bredn<OpAdd>(L, br, q ? T(0) : cabs2(xr));
    const T ar = alpha.re, ai = alpha.im;
    cx<T> tau = {T(0), T(0)}, scal = {T(0), T(0)};
    T beta = ar;
    if (xn2 != T(0) || ai != T(0)) {
        const T nrm = sqrt_t<T>(ar * ar + ai * ai + xn2);
        beta = (ar >= T(0)) ? -nrm : nrm;
        const T ib = fast_rcp(beta);
        tau = cx<T>{(beta - ar) * ib, -ai * ib};
        const T dr_ = ar - beta;
        const T id = fast_rcp(dr_ * dr_ + ai * ai);
        scal = cx<T>{dr_ * id, -ai * id};
    }
    const cx<T> vr = (r == j + 1) ? cx<T>{T(1), T(0)} : (r > j + 1 && r < n) ? cmul(xr, scal) : cx<T>{T(0), T(0)};
    if (q == 0 && r < C) L.V[r] = vr;
    if (threadIdx.x == 0) {
        L.D[j] = L.R[j].re;
        L.E[j] = beta;
        L.TAU[j] = tau;
    }
    __syncthreads();
    cx<T> p = {T(0), T(0)};
#pragma unroll
    for (int k0 = (K0 / (2 * KB)) * (2 * KB); k0 < CH; k0 += 2 * KB) {
#pragma unroll
        for (int k = (k0 > K0 ? k0 : K0); k < (k0 + 2 * KB < CH ? k0 + 2 * KB : CH); ++k)
            cfma(p, h[k], L.V[Q * k + q]);
        asm volatile("" ::: "memory");
    }
    p = cmul(tau, qsum<Q>(p));
    const cx<T> s = bsumn(L, br, q ? cx<T>{T(0), T(0)} : cmulc(vr, p));
    const cx<T> a2 = cscale(cmul(tau, s), T(-0.5));
    const cx<T> wr = (r > j && r < n) ? cadd(p, cmul(a2, vr)) : cx<T>{T(0), T(0)};
    if (q == 0 && r < C) L.W[r] = wr;
    __syncthreads();
#pragma unroll
    for (int k0 = (K0 / (2 * KB)) * (2 * KB); k0 < CH; k0 += 2 * KB) {
#pragma unroll
        for (int k = (k0 > K0 ? k0 : K0); k < (k0 + 2 * KB < CH ? k0 + 2 * KB : CH); ++k) {
            cfms_c(h[k], vr, L.W[Q * k + q]);
            cfms_c(h[k], wr, L.V[Q * k + q]);
        }
        asm volatile("" ::: "memory");
    }
    // park this row's entry of v_j in the column just annihilated (held by
    // the thread q = j % Q, at k = j / Q)
#pragma unroll
    for (int k = (K0 > 0 ? K0 - 1 : 0); k < (K0 + KB < CH ? K0 + KB : CH); ++k)
        if (Q * k + q == j && r > j + 1) h[k] = vr;
    if (r == j + 1) {
#pragma unroll
        for (int k = 0; k < CH; ++k) L.R[Q * k + q] = h[k];
    }
    __syncthreads();
}

template <typename T, int Q, int CH, int RW, int K0>
__device__ __forceinline__ void tri_loop(LRows<T, Q, CH, RW>& L, BRedN& br, cx<T> (&h)[CH], int& j, int n, int r, int q) {
    constexpr int KB = 8 / Q;
    const int jend = (Q * K0 + 7 < n - 1) ? Q * K0 + 7 : n - 1;
    for (; j < jend; ++j) tri_step_share<T, Q, CH, RW, K0>(L, br, h, j, n, r, q);
    if constexpr (K0 + KB < CH) tri_loop<T, Q, CH, RW, K0 + KB>(L, br, h, j, n, r, q);
}

template <typename T, int Q, int CH, int RW>
__device__ __forceinline__ void tridiag_share(LRows<T, Q, CH, RW>& L, BRedN& br, cx<T> (&h)[CH], int n, int r, int q) {
    if (r == 0) {
#pragma unroll
        for (int k = 0; k < CH; ++k) L.R[Q * k + q] = h[k];
    }
    __syncthreads();
    int j = 0;
    tri_loop<T, Q, CH, RW, 0>(L, br, h, j, n, r, q);
    if (threadIdx.x == 0) L.D[n - 1] = L.R[n - 1].re;
    __syncthreads();
}

// Sturm count of tridiag(L.D, L.E2) below sig.
template <typename T, int Q, int CH, int RW>
__device__ __forceinline__ int sturm_below(const LRows<T, Q, CH, RW>& L, int n, T sig, T pivmin) {
    T x = L.D[0] - sig;
    if (fabs(x) < pivmin) x = -pivmin;
    int cnt = x < T(0);
    for (int i = 1; i < n; ++i) {
        x = (L.D[i] - sig) - L.E2[i - 1] * fast_rcp(x);
        if (fabs(x) < pivmin) x = -pivmin;
        cnt += x < T(0);
    }
    return cnt;
}

// ---- eigenvalue k (0-based ascending) of tridiag(D, E): (RW Q + 1)-section
template <typename T, int Q, int CH, int RW>
__device__ __forceinline__ T bisect_share(LRows<T, Q, CH, RW>& L, BRedN& br, int n, int k, int r, int q, T& tnorm,
                                          T& pm) {
    constexpr int NT = RW * Q;   // threads
    T lo = T(3.0e38), hi = T(-3.0e38), em2 = T(0), tn = T(0);
    if (q == 0 && r < n) {
        const T a = (r > 0) ? fabs(L.E[r - 1]) : T(0);
        const T b = (r < n - 1) ? fabs(L.E[r]) : T(0);
        lo = L.D[r] - a - b;
        hi = L.D[r] + a + b;
        tn = fabs(L.D[r]) + a + b;
        if (r < n - 1) {
            L.E2[r] = L.E[r] * L.E[r];
            em2 = L.E2[r];
        }
    }
    lo = bredn<OpMin>(L, br, lo);
    hi = bredn<OpMax>(L, br, hi);
    em2 = bredn<OpMax>(L, br, em2);
    tn = bredn<OpMax>(L, br, tn);
    tnorm = tn;
    const T pivmin = Lim<T>::tiny() * max(T(1), em2);
    pm = pivmin;
    const T pad = T(4) * Lim<T>::eps() * tn + T(2) * pivmin;
    lo -= pad;
    hi += pad;
    // four rounds in FP32 narrow the Gershgorin interval 2.8e8-fold (129
    // sections, the fewest threads) to 4.4e9-fold (257); seven in FP64
    // 7.4e16-fold (257) to 9.4e18-fold (513)
    constexpr int rounds = sizeof(T) == 4 ? 4 : 7;
    const int t = threadIdx.x;
    for (int it = 0; it < rounds; ++it) {
        const T w = (hi - lo) * (T(1) / T(NT + 1));
        const T sig = lo + w * (T)(t + 1);
        const int cnt = sturm_below(L, n, sig, pivmin);
        const int f = bfirstn(L, br, cnt > k);
        const T nhi = (f < NT) ? lo + w * (T)(f + 1) : hi;
        const T nlo = (f > 0) ? lo + w * (T)f : lo;
        lo = nlo;
        hi = nhi;
    }
    return T(0.5) * (lo + hi);
}

// ---- eigenvector of tridiag(D, E) for lam: two rounds of inverse iteration
template <typename T, int Q, int CH, int RW>
__device__ __forceinline__ T tri_eigvec_share(LRows<T, Q, CH, RW>& L, BRedN& br, int n, int r, int q, T lam, T tnorm) {
    constexpr int C = Q * CH;
    const T pert = max(Lim<T>::eps() * tnorm, Lim<T>::tiny());
    if (q == 0 && r < C) L.B[r] = T(0.5) + (T)((r * 37 + 11) & 63) * (T(1) / T(64));
    __syncthreads();
    T xme = T(0);
    T cd = L.D[0] - lam;
    for (int round = 0; round < 2; ++round) {
        T cx_ = L.B[0];
        if (round == 0) {
            T cu = (n > 1) ? L.E[0] : T(0);
            for (int i = 0; i < n - 1; ++i) {
                const T e = L.E[i];
                const T dn = L.D[i + 1] - lam;
                const T un = (i + 1 < n - 1) ? L.E[i + 1] : T(0);
                const T bn = L.B[i + 1];
                const bool keep = fabs(cd) >= fabs(e);
                T pv, qv, sv, xv, ncd, ncu, ncx, f;
                if (keep) {
                    const T cdp = (fabs(cd) < pert) ? (cd < T(0) ? -pert : pert) : cd;
                    f = e * fast_rcp(cdp);
                    pv = cdp; qv = cu; sv = T(0); xv = cx_;
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
                    L.FK[i] = keep ? T(1) : T(0);
                }
                cd = ncd;
                cu = ncu;
                cx_ = ncx;
            }
            if (fabs(cd) < pert) cd = (cd < T(0)) ? -pert : pert;
        } else {
            for (int i = 0; i < n - 1; ++i) {
                const T bn = L.B[i + 1];
                const T f = L.FF[i];
                const bool keep = L.FK[i] != T(0);
                const T xv = keep ? cx_ : bn;
                const T ncx = keep ? bn - f * cx_ : cx_ - f * bn;
                if (threadIdx.x == 0) L.FX[i] = xv;
                cx_ = ncx;
            }
        }
        __syncthreads();
        T x1 = cx_ * fast_rcp(cd);
        T x2 = T(0);
        if (r == n - 1) xme = x1;
        for (int i = n - 2; i >= 0; --i) {
            const T xi = (L.FX[i] - L.FQ[i] * x1 - L.FS[i] * x2) * fast_rcp(L.FP[i]);
            if (r == i) xme = xi;
            x2 = x1;
            x1 = xi;
        }
        const T mx = bredn<OpMax>(L, br, (r < n) ? fabs(xme) : T(0));
        const T sc = (mx > T(0) && isfinite(mx)) ? T(1) / mx : T(1);
        xme *= sc;
        if (q == 0 && r < C) L.B[r] = xme;
        __syncthreads();
    }
    return (r < n) ? xme : T(0);
}

// y <- Q y with the reflectors parked in h (entry r > j+1 of v_j at column j).
template <typename T, int Q, int CH, int RW>
__device__ __forceinline__ cx<T> back_transform_share(LRows<T, Q, CH, RW>& L, BRedN& br, const cx<T> (&h)[CH], int n,
                                                      int r, int q, cx<T> y) {
    constexpr int C = Q * CH;
    // j is a compile-time index into h in every step: two fully unrolled
    // loops (one of 127 steps exceeds the unroller's size limit, and h would
    // then be indexed at run time, in scratch memory)
    auto step = [&](int j) {
        const cx<T> part = (q == j % Q && r > j + 1 && r < n) ? h[j / Q] : cx<T>{T(0), T(0)};
        const cx<T> v = (r == j + 1) ? cx<T>{T(1), T(0)} : qsum<Q>(part);
        const cx<T> s = bsumn(L, br, q ? cx<T>{T(0), T(0)} : cmulc(y, v));
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

// Certification of an FP32 eigenpair (criteria at kCertScale).
template <typename T, int Q, int CH, int RW>
__device__ __forceinline__ bool eig_uncertain_share(LRows<T, Q, CH, RW>& L, BRedN& br, int n, int r, int q, T lam, T tn,
                                                    T pivmin, bool largest, cx<T> y) {
    const T a2 = (r < n) ? cabs2(y) : T(0);
    const T s2 = bredn<OpAdd>(L, br, q ? T(0) : a2);
    const T m2 = bredn<OpMin>(L, br, (r < n) ? a2 : T(3.0e38));
    const T vmin = sqrt(m2 / s2);
    const T delta = T(kCertScale) * Lim<T>::eps() * tn / vmin;
    if (!(delta <= tn)) return true;
    const int c = sturm_below(L, n, largest ? lam - delta : lam + delta, pivmin);
    return largest ? (c <= n - 2) : (c >= 2);
}

template <typename T, int Q, int CH, int RW, typename Tg>
__device__ __forceinline__ bool load_share(LRows<T, Q, CH, RW>& L, BRedN& br, cx<T> (&h)[CH], const cx<Tg>* grow, int n,
                                           int r, int q, bool act, int band) {
    bool fin = true;
#pragma unroll
    for (int k0 = 0; k0 < CH; k0 += 16) {
#pragma unroll
        for (int k = k0; k < (k0 + 16 < CH ? k0 + 16 : CH); ++k) {
            const int c = Q * k + q;
            cx<T> v = {T(0), T(0)};
            if (r < n && c < n) {
                if (act) {
                    const cx<Tg> g = grow[c];
                    v = cx<T>{(T)g.re, (T)g.im};
                } else if (c == r) {
                    v = cx<T>{T(1), T(0)};
                }
                if (band >= 0 && (c - r > band || r - c > band)) v = cx<T>{T(0), T(0)};
            }
            fin = fin && isfinite(v.re) && isfinite(v.im);
            h[k] = v;
        }
        asm volatile("" ::: "memory");
    }
    return bredn<OpMin>(L, br, fin ? 1.0f : 0.0f) > 0.5f;
}

template <typename T, int Q, int CH, int RW, typename Tg>
__device__ __forceinline__ bool hadamard_share(LRows<T, Q, CH, RW>& L, BRedN& br, cx<T> (&h)[CH], const double (&m)[CH],
                                               const cx<Tg>* grow, int n, int r, int q, bool act) {
    bool fin = true;
#pragma unroll
    for (int k0 = 0; k0 < CH; k0 += 16) {
#pragma unroll
        for (int k = k0; k < (k0 + 16 < CH ? k0 + 16 : CH); ++k) {
            const int c = Q * k + q;
            cx<T> v = {T(0), T(0)};
            if (r < n && c < n) {
                cx<T> g = (c == r) ? cx<T>{T(1), T(0)} : cx<T>{T(0), T(0)};
                if (act && c != r) {
                    const cx<Tg> gg = grow[c];
                    g = cx<T>{(T)gg.re, (T)gg.im};
                }
                const T mm = (T)(-m[k]);   // m holds -|Gamma|^-1
                v = cx<T>{mm * g.re, mm * g.im};
                fin = fin && isfinite(v.re + v.im);
            }
            h[k] = v;
        }
        asm volatile("" ::: "memory");
    }
    return bredn<OpMin>(L, br, fin ? 1.0f : 0.0f) > 0.5f;
}

// The estimator chain of one pixel (every thread of the workgroup).  CERT:
// the FP32 result is certified and an uncertified pixel is listed.
template <typename T, int Q, int CH, int RW, int MODE, bool CERT, typename Tg>
__device__ __forceinline__ void solve_rows_pixel(const SolveArgs& a, LRows<T, Q, CH, RW>& L, const cx<Tg>* gam, int np,
                                                 long long pix) {
    constexpr int C = Q * CH;
    BRedN br;
    const int r = threadIdx.x / Q, q = threadIdx.x % Q;
    const int n = a.n;
    const bool act = np >= 0;
    const cx<Tg>* grow = gam + (size_t)min(r, n - 1) * n;

    float code = 0.0f;
    bool failed = false;
    cx<T> h[CH];
    cx<T> y = {T(0), T(0)};
    T tn, pm;
    bool unc = false;
    const T thr_margin = T(kCertThrMargin) * Lim<T>::eps();

    if constexpr (MODE == MODE_MLE) {
        const bool fin = load_share(L, br, h, grow, n, r, q, act, -1);
        tridiag_share(L, br, h, n, r, q);
        const T lmin = bisect_share(L, br, n, 0, r, q, tn, pm);
        if (!fin) { code = -1.0f; failed = true; }
        else if (!(lmin >= T(1.0e-6))) { code = -2.0f; failed = true; }
        if constexpr (CERT) unc = !(lmin >= T(1.0e-6) + thr_margin * tn);
        double m[CH];
        double pmin;
        abs_gamma_share<Q, CH>(m, grow, n, r, q, act, 1.0e-6);
        const bool pd_shift = sweep_share(L, m, n, r, q, pmin);
        if (!failed && !pd_shift) { code = -4.0f; failed = true; }
        if constexpr (CERT) unc = unc || !(pmin >= kCertPivotMin);
        abs_gamma_share<Q, CH>(m, grow, n, r, q, act, 0.0);
        const bool pd = sweep_share(L, m, n, r, q, pmin);
        if (!failed && !pd) { code = -5.0f; failed = true; }
        if constexpr (CERT) unc = unc || !(pmin >= kCertPivotMin);
        const bool hfin = hadamard_share(L, br, h, m, grow, n, r, q, act);
        tridiag_share(L, br, h, n, r, q);
        const T lam = bisect_share(L, br, n, 0, r, q, tn, pm);
        const T yr = tri_eigvec_share(L, br, n, r, q, lam, tn);
        y = back_transform_share(L, br, h, n, r, q, cx<T>{yr, T(0)});
        if (!failed && !hfin) { code = -6.0f; failed = true; }
        if (!failed && !(lam >= T(1.0e-6))) { code = -7.0f; failed = true; }
        if constexpr (CERT) {
            unc = unc || !(lam >= T(1.0e-6) + thr_margin * tn);
            const bool ue = eig_uncertain_share(L, br, n, r, q, lam, tn, pm, false, y);
            unc = fin && (unc || (!failed && ue));
        }
    } else if constexpr (MODE == MODE_EVD) {
        const int band = (a.method == FRINGE_METHOD_STBAS) ? a.bw : -1;
        const bool fin = load_share(L, br, h, grow, n, r, q, act, band);
        tridiag_share(L, br, h, n, r, q);
        const T lam = bisect_share(L, br, n, n - 1, r, q, tn, pm);
        const T yr = tri_eigvec_share(L, br, n, r, q, lam, tn);
        y = back_transform_share(L, br, h, n, r, q, cx<T>{yr, T(0)});
        if (!fin) { code = -6.0f; failed = true; }
        else if (!(lam >= T(1.0e-6))) { code = -7.0f; failed = true; }
        if constexpr (CERT) {
            const bool ue = eig_uncertain_share(L, br, n, r, q, lam, tn, pm, true, y);
            unc = fin && (ue || !(lam >= T(1.0e-6) + thr_margin * tn));
        }
    } else {
        double m[CH];
        double pmin;
        bool fin = abs_gamma_share<Q, CH>(m, grow, n, r, q, act, 0.0);
        fin = bredn<OpMin>(L, br, fin ? 1.0f : 0.0f) > 0.5f;
        const bool pd = sweep_share(L, m, n, r, q, pmin);
        const bool hfin = hadamard_share(L, br, h, m, grow, n, r, q, act);
        const bool run_evd = !pd || !hfin;   // uniform over the workgroup
        if constexpr (CERT) unc = !(pmin >= kCertPivotMin);
        if (!run_evd) {   // the EMI eigenpair; a pixel that falls back to EVD skips it
            tridiag_share(L, br, h, n, r, q);
            const T lam = bisect_share(L, br, n, 0, r, q, tn, pm);
            const T yr = tri_eigvec_share(L, br, n, r, q, lam, tn);
            y = back_transform_share(L, br, h, n, r, q, cx<T>{yr, T(0)});
            if constexpr (CERT) unc = unc || eig_uncertain_share(L, br, n, r, q, lam, tn, pm, false, y);
        }
        if (act && fin && run_evd) {   // uniform: the whole workgroup is one pixel
            load_share(L, br, h, grow, n, r, q, act, -1);
            tridiag_share(L, br, h, n, r, q);
            const T le = bisect_share(L, br, n, n - 1, r, q, tn, pm);
            const T ye = tri_eigvec_share(L, br, n, r, q, le, tn);
            y = back_transform_share(L, br, h, n, r, q, cx<T>{ye, T(0)});
            if constexpr (CERT) unc = unc || eig_uncertain_share(L, br, n, r, q, le, tn, pm, true, y);
        }
        if (!fin) { code = -1.0f; failed = true; }
        unc = unc && fin;
    }

    // ---- referenced phases (evd.cpp:738-749) ----
    if (q == 0 && r == a.ref) L.YREF = y;
    __syncthreads();
    const cx<T> vref = L.YREF;
    bool okr, oku, okv;
    const cx<T> cref = unit_or(cconj(vref), cx<T>{T(1), T(0)}, okr);
    const cx<T> u = unit_or(cmulc(y, vref), cx<T>{T(0), T(0)}, oku);
    const cx<T> ot = oku ? u : (okr ? cref : unit_or(y, cx<T>{T(1), T(0)}, okv));
    cx<float> o = {(float)ot.re, (float)ot.im};
    if (r == a.ref) o = cx<float>{1.0f, 0.0f};
    if (q == 0 && r < C) L.OUT[r] = o;
    __syncthreads();

    const bool write_ok = act && !failed;
    // ---- compressed SLC (evd.cpp:755-762) ----
    cx<double> cs = {0.0, 0.0};
    if (q == 0 && r >= a.ref && r < n && write_ok) {
        const float2 z = a.slc[(long long)r * a.npx + pix];
        cs.re = (double)(z.x * o.re + z.y * o.im);
        cs.im = (double)(z.y * o.re - z.x * o.im);
    }
    cs.re = bredn<OpAdd>(L, br, cs.re);
    cs.im = bredn<OpAdd>(L, br, cs.im);
    // ---- temporal coherence (evd.cpp:770-786), pairs dealt out circulantly ----
    const bool stbas = (a.method == FRINGE_METHOD_STBAS);
    cx<float> tc = {0.0f, 0.0f};
    if (q == 0 && r < n && write_ok) {
        bool okx;
        const cx<float> ci = unit_or(cx<float>{o.re, -o.im}, cx<float>{1.0f, 0.0f}, okx);
        const int half = (n - 1) / 2;
        const int dmax = ((n & 1) == 0 && r < n / 2) ? half + 1 : half;
        int jd = r;
        for (int d = 1; d <= dmax; ++d) {
            jd = jd + 1 == n ? 0 : jd + 1;
            const int lo = min(r, jd), hi = max(r, jd);
            if (stbas && hi - lo > a.bw) continue;
            const cx<Tg> gv = grow[jd];
            const cx<float> ph = unit_or(cx<float>{(float)gv.re, (float)gv.im}, cx<float>{1.0f, 0.0f}, okx);
            const cx<float> oj = L.OUT[jd];
            cx<float> t = cmul(cmul(ph, ci), oj);
            if (r > jd) t = cconj(t);
            tc = cadd(tc, t);
        }
    }
    tc = bsumn(L, br, tc);

    if (q == 0 && r < n)
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
        if constexpr (CERT) {
            if (act && unc) {
                const int e = atomicAdd(a.cert_count, 1);
                if (e < a.cert_cap) a.cert_list[e] = (int)pix;
            }
        }
    }
}

// Least waves per SIMD the register budget is held to: FP32 shares of up to
// 48 columns three (a few registers spilled at 48), of 56 and 64 two; FP64
// two.  The smaller shares take fewer registers and reach more (up to four).
template <typename T, int CH> constexpr int rows_waves_per_eu() { return sizeof(T) == 4 && CH <= 48 ? 3 : 2; }

// One pixel per workgroup of RW Q threads; slot s of the round is pixel
// first + s, its coherence complex Tg.
template <typename T, int Q, int CH, int RW, int MODE, typename Tg>
__global__ __launch_bounds__(RW * Q) __attribute__((amdgpu_waves_per_eu(rows_waves_per_eu<T, CH>())))
void solve_rows_kernel(SolveArgs a) {
    __shared__ LRows<T, Q, CH, RW> L;
    const int slot = (int)(gridDim.x - 1 - blockIdx.x);   // last-first, as the other solvers
    if (slot >= a.slots) return;                            // the whole workgroup
    const cx<Tg>* gam = (const cx<Tg>*)a.gamma + (size_t)slot * a.n * a.n;
    solve_rows_pixel<T, Q, CH, RW, MODE, sizeof(T) == 4, Tg>(a, L, gam, a.npix[slot], a.first + slot);
}

// FP64 re-solve of the listed pixels: each workgroup loops over its share of
// the list.  The waves compute the listed pixel's FP64 coherence with the
// plain path's Gram (gram_f64_pixel), wave w of NW the column blocks w,
// w + NW, ..., each synchronising only itself (with the whole workgroup
// taking part in every block, its two barriers per position over eight waves
// cost more than the chain), into the workgroup's global area; then the
// whole workgroup runs the FP64 chain on it.
// Rows per lane of that Gram (a wave's 64 lanes hold rows g + 64 q, q < R:
// one up to 64 dates, which leaves the 64-row kernels without spilled
// registers) and its column block (each entry's sum is the same whatever
// the rows and block; 8 instead of 16 above 64 dates spilled no less and
// was no faster).
template <int RW> constexpr int resolve_gram_rows() { return RW > 64 ? 2 : 1; }
constexpr int kResolveRowsCB = 16;

template <int Q, int CH, int RW, int MODE>
__global__ __launch_bounds__(RW * Q) __attribute__((amdgpu_waves_per_eu(rows_waves_per_eu<double, CH>())))
void resolve_rows_kernel(SolveArgs a) {
    __shared__ LRows<double, Q, CH, RW> L;
    extern __shared__ __align__(16) char smem[];
    const int n = a.n;
    cx<double>* gam = (cx<double>*)((char*)a.awork + a.awork_stride * (size_t)blockIdx.x);
    const int cnt = min(*a.list_count, a.list_cap);
    for (int e = blockIdx.x; e < cnt; e += gridDim.x) {
        const long long pix = a.list[e];
        const int w = threadIdx.x >> 6;
        const int np = gram_f64_pixel<64, resolve_gram_rows<RW>(), kResolveRowsCB, true>(
            a.g64, true, pix, gam, smem + w * gram_f64_group_bytes(n, a.g64.Nx, a.g64.Ny, kResolveRowsCB),
                                                       w, LRows<double, Q, CH, RW>::NW);
        __threadfence();
        __syncthreads();
        solve_rows_pixel<double, Q, CH, RW, MODE, false, double>(a, L, gam, np, pix);
        __syncthreads();
    }
}

// Launchers (rows32.hip: T = float, Q = 2; rows64.hip: T = double, Q = 4),
// Q CH columns per row and RW rows, the fewest that hold n.
void launch_rows_solver_f32(const SolveArgs& a, hipStream_t s);
void launch_rows_solver_f64(const SolveArgs& a, hipStream_t s);
// a.awork / a.awork_stride: the per-workgroup areas (at least n^2 complex
// doubles each), blocks of them; a.g64, a.list*: the FP64 Gram and the list.
void launch_resolve_rows(const SolveArgs& a, unsigned blocks, hipStream_t s);
// LDS of the re-solve's workgroup (static chain state plus one FP64 Gram
// scratch per wave), for the launch checks.
size_t resolve_rows_lds_bytes(int n, int Nx, int Ny);

}  // namespace fringe
