// solve32.hpp -- register-resident per-pixel estimator for n <= 64 dates.
//
// NR = 32 (n <= 32): a wave holds two pixels and lane r of a 32-lane half
// owns row r of every matrix of its pixel in registers.  NR = 64
// (32 < n <= 64): one pixel per wave, lane r owns row r.  LDS (about 3.5 KB per pixel) is used only
// to publish one row or one vector to the other lanes, so many waves fit on
// a CU and the arithmetic is not serialised behind LDS round trips.
//
// Stages (same semantics as the generic LDS solver in link_solve.hpp):
//   * inverse of the real |Gamma| by the symmetric sweep operator in FP64,
//     with the matrix in 4 x 8 blocks (lane (a, b) of a half holds rows
//     4a..4a+3, columns 8b..8b+7): per pivot a lane reads 4 multipliers and
//     8 pivot-row entries from LDS instead of a whole row, then the blocks go
//     back to one row per lane a block column at a time;
//   * H = |Gamma|^-1 o Gamma (or Gamma itself for EVD) into registers;
//   * Householder tridiagonalisation: the owner of row j publishes it, the
//     reflector v_j = conj(row j) * scal_j is applied, and every lane parks
//     its own entry of v_j in the register of the column just annihilated,
//     where the back-transformation finds it;
//   * eigenvalue by 33-section of Sturm counts, eigenvector of T by inverse
//     iteration (every lane runs the same pivoted elimination chain on
//     broadcast operands), back-transformation by the stored reflectors;
//   * referenced phases, compressed SLC and temporal coherence.
#pragma once
#include "engine_internal.hpp"
#include "link_solve.hpp"

namespace fringe {

template <typename T, int NR> struct LdsR {
    cx<T> R[NR];      // published row
    cx<T> V[NR];
    cx<T> W[NR];
    cx<T> TAU[NR];
    cx<T> YREF;
    T D[NR], E[NR], E2[NR];
    T B[NR];          // inverse-iteration right-hand side
    T FP[NR], FQ[NR], FS[NR], FX[NR];   // eliminated rows
    T FF[NR], FK[NR];  // multiplier of elimination step i, 1 where its pivot was kept
    double TR[NR];    // sweep: published pivot row
    double PV;        // sweep: pivot
    double XT[NR / 8][NR];   // sweep: one block column on its way to the row layout
    cx<float> OUT[NR];
};

// Blocks of the FP64 sweep: 8 columns wide, NR / 8 rows high; lane r holds
// row block r / (NR / 8) and column block r % (NR / 8).
template <int NR> struct Blk {
    static constexpr int BC = 8, NCB = NR / 8, BR = NR / 8;
};

__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T> __device__ __forceinline__ T fast_rcp(T x) { return T(1) / x; }
template <> __device__ __forceinline__ float fast_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
// 1/x in double: hardware estimate refined by two Newton steps (~1 ulp).
template <> __device__ __forceinline__ double fast_rcp<double>(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}

// ---- reductions over the 32 lanes of a half-wave without LDS --------------
// quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror and row_mirror
// combine 16 lanes through DPP; v_permlane16_swap then exchanges the two
// 16-lane rows of the half.  Every lane ends with the same, bitwise
// identical value (each stage adds two equal-rank partials commutatively).
template <int CTRL> __device__ __forceinline__ uint32_t dpp32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
struct OpAdd { template <typename T> __device__ static T f(T a, T b) { return a + b; } };
struct OpMax { template <typename T> __device__ static T f(T a, T b) { return max(a, b); } };
struct OpMin { template <typename T> __device__ static T f(T a, T b) { return min(a, b); } };

template <typename Op> __device__ __forceinline__ float red32(float x) {
    x = Op::f(x, __uint_as_float(dpp32<0xB1>(__float_as_uint(x))));
    x = Op::f(x, __uint_as_float(dpp32<0x4E>(__float_as_uint(x))));
    x = Op::f(x, __uint_as_float(dpp32<0x141>(__float_as_uint(x))));
    x = Op::f(x, __uint_as_float(dpp32<0x140>(__float_as_uint(x))));
    const auto s = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    return Op::f(__uint_as_float(s[0]), __uint_as_float(s[1]));
}
template <int CTRL> __device__ __forceinline__ double dpp64(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = dpp32<CTRL>((uint32_t)u), hi = dpp32<CTRL>((uint32_t)(u >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
template <typename Op> __device__ __forceinline__ double red32(double x) {
    x = Op::f(x, dpp64<0xB1>(x));
    x = Op::f(x, dpp64<0x4E>(x));
    x = Op::f(x, dpp64<0x141>(x));
    x = Op::f(x, dpp64<0x140>(x));
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)u, (uint32_t)u, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(u >> 32), (uint32_t)(u >> 32), false, false);
    const double a = __builtin_bit_cast(double, ((uint64_t)hi[0] << 32) | lo[0]);
    const double b = __builtin_bit_cast(double, ((uint64_t)hi[1] << 32) | lo[1]);
    return Op::f(a, b);
}
// over the NR lanes of a pixel: for NR = 64 the two half-wave results are
// combined once more through v_permlane32_swap
template <typename Op, int NR> __device__ __forceinline__ float redR(float x) {
    x = red32<Op>(x);
    if constexpr (NR == 64) {
        const auto s = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
        x = Op::f(__uint_as_float(s[0]), __uint_as_float(s[1]));
    }
    return x;
}
template <typename Op, int NR> __device__ __forceinline__ double redR(double x) {
    x = red32<Op>(x);
    if constexpr (NR == 64) {
        const uint64_t u = __builtin_bit_cast(uint64_t, x);
        const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)u, (uint32_t)u, false, false);
        const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(u >> 32), (uint32_t)(u >> 32), false, false);
        const double a = __builtin_bit_cast(double, ((uint64_t)hi[0] << 32) | lo[0]);
        const double b = __builtin_bit_cast(double, ((uint64_t)hi[1] << 32) | lo[1]);
        x = Op::f(a, b);
    }
    return x;
}
template <int NR, typename T> __device__ __forceinline__ T sumR(T x) { return redR<OpAdd, NR>(x); }
template <int NR, typename T> __device__ __forceinline__ T maxR(T x) { return redR<OpMax, NR>(x); }
template <int NR, typename T> __device__ __forceinline__ T minR(T x) { return redR<OpMin, NR>(x); }
template <int NR, typename T> __device__ __forceinline__ cx<T> sumR(cx<T> x) { return cx<T>{sumR<NR>(x.re), sumR<NR>(x.im)}; }

// ---- FP64 inverse of |Gamma| - shift*I in BR x 8 blocks -------------------
// Lane (a, b) = (r / NCB, r % NCB) holds M[i][j] = A(BR a + i, 8b + j)
// (BR = 4, NCB = 4 for NR = 32; BR = 8, NCB = 8 for NR = 64).
// Rows and columns >= n hold the identity (never swept, never mixed in).
template <int NR, typename Tg>
__device__ __forceinline__ bool abs_gamma_block(double (&M)[Blk<NR>::BR][8], const cx<Tg>* gam, int n, int r, bool act,
                                       double shift) {
    constexpr int BR = Blk<NR>::BR, NCB = Blk<NR>::NCB;
    const int a = r / NCB, b = r % NCB;
    bool fin = true;
#pragma unroll
    for (int i = 0; i < BR; ++i) {
        const int row = BR * a + i;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int col = 8 * b + j;
            double v = row == col ? 1.0 - shift : 0.0;
            if (act && row < n && col < n && row != col) {
                const cx<Tg> g = gam[(size_t)row * n + col];
                if constexpr (sizeof(Tg) == 4)   // float Gamma: one v_sqrt_f32 (~1 ulp)
                    v = (double)__builtin_amdgcn_sqrtf(g.re * g.re + g.im * g.im);
                else
                    v = sqrt(g.re * g.re + g.im * g.im);
                fin = fin && isfinite(v);
            }
            M[i][j] = v;
        }
    }
    return fin;
}

// Symmetric sweep operator: sweeping pivot k maps
//   A_kk -> -1/A_kk,  A_rk -> A_rk/A_kk,  A_kc -> A_kc/A_kk,
//   A_rc -> A_rc - A_rk A_kc / A_kk                        (r, c != k)
// and keeps A symmetric, so the multiplier of row r is entry r of the
// published pivot row P.  With P_k = A_kk - 1 every entry but (k, k) is
// A_rc - (P_r / p) P_c; (k, k) is set to -1/p.  After pivots 0..n-1 the
// blocks hold -A^-1.  Returns whether every pivot was positive (A is
// positive definite).
template <typename T, int NR>
__device__ __forceinline__ bool sweep_block(LdsR<T, NR>& L, double (&M)[Blk<NR>::BR][8], int n, int r,
                                            double& pmin) {
    constexpr int BR = Blk<NR>::BR, NCB = Blk<NR>::NCB;
    const int a = r / NCB, b = r % NCB;
    bool ok = true;
    pmin = 1.0;
    for (int q = 0; 8 * q < n; ++q) {
#pragma unroll
        for (int kj = 0; kj < 8; ++kj) {
            const int k = 8 * q + kj;
            if (k >= n) break;
            const int ki = kj % BR;   // row k is row ki of the blocks with a == k / BR
            if (a == k / BR) {
                double pr[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) pr[j] = M[ki][j];
                if (b == q) {   // column k is column kj of the blocks with b == q
                    L.PV = pr[kj];
                    pr[kj] = pr[kj] - 1.0;
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) L.TR[8 * b + j] = pr[j];
            }
            wsync();
            const double p = L.PV;
            if (ok) pmin = fmin(pmin, fabs(p));   // up to the first pivot that is not positive
            ok = ok && (p > 0.0);
            const double piv = fast_rcp(p);
            double pc[8], g[BR];
#pragma unroll
            for (int j = 0; j < 8; ++j) pc[j] = L.TR[8 * b + j];
#pragma unroll
            for (int i = 0; i < BR; ++i) g[i] = L.TR[BR * a + i] * piv;
#pragma unroll
            for (int i = 0; i < BR; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) M[i][j] = fma(-g[i], pc[j], M[i][j]);
            if (a == k / BR && b == q) M[ki][kj] = -piv;
            wsync();
        }
    }
    return ok;
}

// Blocks -> rows: m[c] = M(r, c).  Block column j (columns j, 8 + j, 16 + j,
// ...) passes through LDS at a time, so the registers it frees are the ones
// the row layout fills.
template <typename T, int NR>
__device__ __forceinline__ void blocks_to_rows(LdsR<T, NR>& L, const double (&M)[Blk<NR>::BR][8], double (&m)[NR], int r) {
    constexpr int BR = Blk<NR>::BR, NCB = Blk<NR>::NCB;
    const int a = r / NCB, b = r % NCB;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
#pragma unroll
        for (int i = 0; i < BR; ++i) L.XT[b][BR * a + i] = M[i][j];
        wsync();
#pragma unroll
        for (int k = 0; k < NCB; ++k) m[8 * k + j] = L.XT[k][r];
        wsync();
    }
}

// ---- Householder tridiagonalisation of the rows in h ---------------------
// Step j of the reduction, for C0 - 1 <= j <= C0 + STEP - 2: columns below
// C0 hold only finished entries and parked reflectors, so they are skipped.
// Columns from NC on (n <= NC <= NR) are skipped too: there h, V and W hold
// +0 in every step, so each skipped product is a zero added to a sum that
// is never -0 (it starts at +0, and an FMA returns -0 only from a -0
// addend), and each skipped update leaves a +0 that nothing reads: the
// results are the same bits as with NC = NR.
template <typename T, int NR, int C0, int STEP, int NC = NR>
__device__ __forceinline__ void tri_step(LdsR<T, NR>& L, cx<T> (&h)[NR], int j, int n, int r) {
    // row j was published at the end of the previous step (or before the first)
    const cx<T> alpha = cconj(L.R[j + 1]);
    const cx<T> xr = (r > j + 1 && r < n) ? cconj(L.R[r]) : cx<T>{T(0), T(0)};
    const T xn2 = sumR<NR>(cabs2(xr));
    const T ar = alpha.re, ai = alpha.im;
    cx<T> tau = {T(0), T(0)}, scal = {T(0), T(0)};
    T beta = ar;
    if (xn2 != T(0) || ai != T(0)) {
        T nrm;
        if constexpr (sizeof(T) == 4)   // one v_sqrt_f32 (~1 ulp): the reflector stays unitary to rounding
            nrm = __builtin_amdgcn_sqrtf(ar * ar + ai * ai + xn2);
        else
            nrm = sqrt(ar * ar + ai * ai + xn2);
        beta = (ar >= T(0)) ? -nrm : nrm;
        const T ib = fast_rcp(beta);
        tau = cx<T>{(beta - ar) * ib, -ai * ib};
        const T dr_ = ar - beta;
        const T id = fast_rcp(dr_ * dr_ + ai * ai);
        scal = cx<T>{dr_ * id, -ai * id};
    }
    const cx<T> vr = (r == j + 1) ? cx<T>{T(1), T(0)} : (r > j + 1 && r < n) ? cmul(xr, scal) : cx<T>{T(0), T(0)};
    L.V[r] = vr;
    if (r == 0) {
        L.D[j] = L.R[j].re;
        L.E[j] = beta;
        L.TAU[j] = tau;
    }
    wsync();
    cx<T> p = {T(0), T(0)};
#pragma unroll
    for (int c = C0; c < NC; ++c) cfma(p, h[c], L.V[c]);
    p = cmul(tau, p);
    const cx<T> s = sumR<NR>(cmulc(vr, p));
    const cx<T> a2 = cscale(cmul(tau, s), T(-0.5));
    const cx<T> wr = (r > j && r < n) ? cadd(p, cmul(a2, vr)) : cx<T>{T(0), T(0)};
    L.W[r] = wr;
    wsync();
#pragma unroll
    for (int c = C0; c < NC; ++c) {
        cfms_c(h[c], vr, L.W[c]);
        cfms_c(h[c], wr, L.V[c]);
    }
    // column j of the rows below j+1 has just been annihilated: park this
    // lane's entry of v_j there for the back-transformation (no later step
    // reads it: V and W vanish at column j from here on)
#pragma unroll
    for (int c = (C0 > 0 ? C0 - 1 : 0); c < (C0 + STEP - 1 < NC ? C0 + STEP - 1 : NC); ++c)
        if (c == j && r > j + 1) h[c] = vr;
    // publish row j + 1 for the next step (everyone is past reading row j;
    // entries from NC on are never read)
    if (r == j + 1) {
#pragma unroll
        for (int c = 0; c < NC; ++c) L.R[c] = h[c];
    }
    wsync();
}

// Steps C0 - 1 .. C0 + STEP - 2 (the last block: up to n - 2), then the
// next block of STEP columns: one loop per quad (NR = 32) or octet (NR =
// 64) of columns left to update (separate loops keep the register
// allocation of each variant to itself).
template <typename T, int NR, int NC, int C0, int STEP>
__device__ __forceinline__ void tri_blocks(LdsR<T, NR>& L, cx<T> (&h)[NR], int& j, int n, int r) {
    if constexpr (C0 + STEP < NC) {
        for (; j < min(C0 + STEP - 1, n - 1); ++j) tri_step<T, NR, C0, STEP, NC>(L, h, j, n, r);
        tri_blocks<T, NR, NC, C0 + STEP, STEP>(L, h, j, n, r);
    } else {
        for (; j < n - 1; ++j) tri_step<T, NR, C0, STEP, NC>(L, h, j, n, r);
    }
}

// NC: the columns processed (n <= NC <= NR; the last block may be partial).
template <typename T, int NR, int NC = NR>
__device__ __forceinline__ void tridiag(LdsR<T, NR>& L, cx<T> (&h)[NR], int n, int r) {
    static_assert(NC > 0 && NC <= NR, "columns of the lanes' rows");
    if (r == 0) {
#pragma unroll
        for (int c = 0; c < NC; ++c) L.R[c] = h[c];
    }
    wsync();
    // columns below j+1 carry zero reflector entries
    int j = 0;
    tri_blocks<T, NR, NC, 0, NR == 32 ? 4 : 8>(L, h, j, n, r);
    // last diagonal entry: row n - 1 is the one published last
    if (r == 0) L.D[n - 1] = L.R[n - 1].re;
    wsync();
}

// ---- eigenvalue k (0-based ascending) of tridiag(D, E) --------------------
template <typename T, int NR>
__device__ __forceinline__ T bisect(LdsR<T, NR>& L, int n, int k, int r, T& tnorm, T& pm) {
    T lo = T(3.0e38), hi = T(-3.0e38), em2 = T(0), tn = T(0);
    if (r < n) {
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
    lo = minR<NR>(lo);
    hi = maxR<NR>(hi);
    em2 = maxR<NR>(em2);
    tn = maxR<NR>(tn);
    wsync();
    tnorm = tn;
    const T pivmin = Lim<T>::tiny() * max(T(1), em2);
    pm = pivmin;
    const T pad = T(4) * Lim<T>::eps() * tn + T(2) * pivmin;
    lo -= pad;
    hi += pad;
    // (NR+1)-section shrinks the Gershgorin interval (NR+1)-fold per round:
    // 5 rounds of 33 or 4 of 65 reach ~1e-7 |T| (enough for the
    // inverse-iteration shift in FP32), 10 of 33 or 9 of 65 FP64 resolution
    const int rounds = sizeof(T) == 4 ? (NR == 32 ? 5 : 4) : (NR == 32 ? 10 : 9);
    for (int it = 0; it < rounds; ++it) {
        const T w33 = (hi - lo) * T(1.0 / (NR + 1));   // the same expression gives every section point
        const T sig = lo + w33 * T(r + 1);
        T q = L.D[0] - sig;
        if (fabs(q) < pivmin) q = -pivmin;
        int cnt = q < T(0);
        for (int i = 1; i < n; ++i) {
            q = (L.D[i] - sig) - L.E2[i - 1] * fast_rcp(q);
            if (fabs(q) < pivmin) q = -pivmin;
            cnt += q < T(0);
        }
        const uint64_t b = group_ballot<NR>(cnt > k);
        const int f = b ? (int)__builtin_ctzll(b) : NR;
        const T nhi = (f < NR) ? lo + w33 * T(f + 1) : hi;
        const T nlo = (f > 0) ? lo + w33 * T(f) : lo;
        lo = nlo;
        hi = nhi;
    }
    return T(0.5) * (lo + hi);
}

// ---- certification of an FP32 eigenpair (criteria in link_solve.hpp) ------
// Number of eigenvalues of tridiag(D, E) below sig (the Sturm count of
// bisect, evaluated by every lane alike).
template <typename T, int NR>
__device__ __forceinline__ int sturm_below(const LdsR<T, NR>& L, int n, T sig, T pivmin) {
    T q = L.D[0] - sig;
    if (fabs(q) < pivmin) q = -pivmin;
    int cnt = q < T(0);
    for (int i = 1; i < n; ++i) {
        q = (L.D[i] - sig) - L.E2[i - 1] * fast_rcp(q);
        if (fabs(q) < pivmin) q = -pivmin;
        cnt += q < T(0);
    }
    return cnt;
}

// True when the eigenvector y (lane r holds component r) of eigenvalue lam
// of tridiag(D, E) (norm bound tn) is not certified: another eigenvalue lies
// within kCertScale eps |T| / min_t |v_t| of lam, or v has a zero or
// non-finite component.  `largest`: lam is the largest eigenvalue, else the
// smallest.
template <typename T, int NR>
__device__ __forceinline__ bool eig_uncertain(const LdsR<T, NR>& L, int n, int r, T lam, T tn, T pivmin,
                                              bool largest, cx<T> y) {
    const T a2 = (r < n) ? cabs2(y) : T(0);
    const T s2 = sumR<NR>(a2);
    const T m2 = minR<NR>((r < n) ? a2 : T(3.0e38));
    const T vmin = sqrt(m2 / s2);
    const T delta = T(kCertScale) * Lim<T>::eps() * tn / vmin;
    if (!(delta <= tn)) return true;
    const int c = sturm_below(L, n, largest ? lam - delta : lam + delta, pivmin);
    return largest ? (c <= n - 2) : (c >= 2);
}

// ---- eigenvector of tridiag(D, E) for lam: 2 rounds of inverse iteration --
// Every lane of the half runs the same elimination chain on broadcast
// operands; lane 0 records the eliminated rows for the back substitution,
// and lane i keeps component i of the result.  The elimination of
// T - lam I (pivot choices, multipliers, eliminated rows) is the same in
// both rounds, so the second round only carries the new right-hand side
// through the recorded steps.
template <typename T, int NR>
__device__ __forceinline__ T tri_eigvec(LdsR<T, NR>& L, int n, int r, T lam, T tnorm) {
    const T pert = max(Lim<T>::eps() * tnorm, Lim<T>::tiny());
    L.B[r] = T(0.5) + T((r * 37 + 11) & 63) * T(1.0 / 64.0);
    wsync();
    T xme = T(0);
    T cd = L.D[0] - lam;   // ends as the last pivot, reused by the second pass
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
                if (r == 0) {
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
                if (r == 0) L.FX[i] = xv;
                cx_ = ncx;
            }
        }
        wsync();
        T x1 = cx_ * fast_rcp(cd);   // x_{n-1}
        T x2 = T(0);                 // x_{i+2}
        if (r == n - 1) xme = x1;
        for (int i = n - 2; i >= 0; --i) {
            const T xi = (L.FX[i] - L.FQ[i] * x1 - L.FS[i] * x2) * fast_rcp(L.FP[i]);
            if (r == i) xme = xi;
            x2 = x1;
            x1 = xi;
        }
        T mx = maxR<NR>((r < n) ? fabs(xme) : T(0));
        const T sc = (mx > T(0) && isfinite(mx)) ? T(1) / mx : T(1);
        xme *= sc;
        wsync();
        L.B[r] = xme;
        wsync();
    }
    return (r < n) ? xme : T(0);
}

// y <- Q y with Q = H_0 ... H_{n-2}; entry r > j+1 of v_j is h[j] of lane r
// (unrolled over j so the register index is static).
template <typename T, int NR>
__device__ __forceinline__ cx<T> back_transform(LdsR<T, NR>& L, const cx<T> (&h)[NR], int n, int r, cx<T> y) {
#pragma unroll
    for (int j = NR - 2; j >= 0; --j) {
        if (j > n - 2) continue;   // uniform
        const cx<T> v = (r == j + 1) ? cx<T>{T(1), T(0)} : (r > j + 1 && r < n) ? h[j] : cx<T>{T(0), T(0)};
        const cx<T> s = sumR<NR>(cmulc(y, v));
        y = csub(y, cmul(v, cmul(L.TAU[j], s)));
    }
    return y;
}

template <typename T, typename Tg, int NR>
__device__ __forceinline__ bool load_rows(cx<T> (&h)[NR], const cx<Tg>* grow, int n, int r, bool act, int band) {
    bool fin = true;
#pragma unroll
    for (int c = 0; c < NR; ++c) {
        cx<Tg> v = {Tg(0), Tg(0)};
        if (r < n && c < n) {
            if (act) v = grow[c];
            else if (c == r) v = cx<Tg>{Tg(1), Tg(0)};
            if (band >= 0 && (c - r > band || r - c > band)) v = cx<Tg>{Tg(0), Tg(0)};
        }
        fin = fin && isfinite(v.re) && isfinite(v.im);
        h[c] = cx<T>{(T)v.re, (T)v.im};
    }
    return minR<NR>(fin ? 1.0f : 0.0f) > 0.5f;
}

template <typename T, typename Tg, int NR>
__device__ __forceinline__ bool hadamard_rows(cx<T> (&h)[NR], const double (&m)[NR], const cx<Tg>* grow, int n,
                                     int r, bool act) {
    bool fin = true;
#pragma unroll
    for (int c = 0; c < NR; ++c) {
        cx<T> v = {T(0), T(0)};
        if (r < n && c < n) {
            cx<Tg> g = (c == r) ? cx<Tg>{Tg(1), Tg(0)} : cx<Tg>{Tg(0), Tg(0)};
            if (act && c != r) g = grow[c];
            if constexpr (sizeof(T) == 4) {
                // FP32 solver: the inverse entry rounded to float, one product
                // each; an infinite or NaN part makes the sum non-finite
                const float mm = (float)(-m[c]);   // m holds -|Gamma|^-1
                v = cx<T>{mm * (float)g.re, mm * (float)g.im};
                fin = fin && isfinite(v.re + v.im);
            } else {
                const double mm = -m[c];
                const double hr = mm * (double)g.re, hi = mm * (double)g.im;
                fin = fin && isfinite(hr) && isfinite(hi);
                v = cx<T>{(T)hr, (T)hi};
            }
        }
        h[c] = v;
    }
    return minR<NR>(fin ? 1.0f : 0.0f) > 0.5f;
}

enum { MODE_EMI = 0, MODE_EVD = 1, MODE_MLE = 2 };

// MODE_EMI: phase_link (EMI with EVD fallback, phase_link.cpp:540-618)
// MODE_EVD: evd EVD / STBAS (evd.cpp:689-732)
// MODE_MLE: evd MLE with its failure codes (evd.cpp:584-688)
// One pixel's estimator chain on the NR lanes of its group: gam is its
// coherence (global memory, or LDS in the FP64 re-solve), np its neighbour
// count (-1: not linked).  CERT: the FP32 result is checked and, when it is
// not certified, the pixel is appended to a.cert_list for the FP64 re-solve.
template <typename T, typename Tg, int MODE, int NR, bool CERT, int NC = NR>
__device__ __forceinline__ void solve_reg_pixel(const SolveArgs& a, LdsR<T, NR>& L, int r, bool inslot, int np,
                                                const cx<Tg>* gam, long long pix) {
    const int n = a.n;
    const bool act = np >= 0;
    const cx<Tg>* grow = gam + (size_t)min(r, n - 1) * n;

    float code = 0.0f;
    bool failed = false;
    cx<T> h[NR];
    cx<T> y = {T(0), T(0)};
    T tn, pm;
    bool unc = false;   // CERT: the FP32 result is not certified

    if constexpr (MODE == MODE_MLE) {
        const bool fin = load_rows<T, Tg, NR>(h, grow, n, r, act, -1);
        tridiag<T, NR, NC>(L, h, n, r);
        const T lmin = bisect<T, NR>(L, n, 0, r, tn, pm);
        if (!fin) { code = -1.0f; failed = true; }
        else if (!(lmin >= T(1.0e-6))) { code = -2.0f; failed = true; }
        if constexpr (CERT) unc = !(lmin >= T(1.0e-6) + T(kCertThrMargin) * Lim<T>::eps() * tn);
        double M[Blk<NR>::BR][8], m[NR];
        double pmin1, pmin2;
        abs_gamma_block<NR, Tg>(M, gam, n, r, act, 1.0e-6);
        const bool pd_shift = sweep_block<T, NR>(L, M, n, r, pmin1);
        if (!failed && !pd_shift) { code = -4.0f; failed = true; }
        if constexpr (CERT) unc = unc || !(pmin1 >= kCertPivotMin);
        abs_gamma_block<NR, Tg>(M, gam, n, r, act, 0.0);
        const bool pd = sweep_block<T, NR>(L, M, n, r, pmin2);
        if (!failed && !pd) { code = -5.0f; failed = true; }
        if constexpr (CERT) unc = unc || !(pmin2 >= kCertPivotMin);
        blocks_to_rows<T, NR>(L, M, m, r);
        const bool hfin = hadamard_rows<T, Tg, NR>(h, m, grow, n, r, act);
        tridiag<T, NR, NC>(L, h, n, r);
        const T lam = bisect<T, NR>(L, n, 0, r, tn, pm);
        const T yr = tri_eigvec<T, NR>(L, n, r, lam, tn);
        y = back_transform<T, NR>(L, h, n, r, cx<T>{yr, T(0)});
        if (!failed && !hfin) { code = -6.0f; failed = true; }
        if (!failed && !(lam >= T(1.0e-6))) { code = -7.0f; failed = true; }
        if constexpr (CERT) {
            unc = unc || !(lam >= T(1.0e-6) + T(kCertThrMargin) * Lim<T>::eps() * tn);
            const bool ue = eig_uncertain<T, NR>(L, n, r, lam, tn, pm, false, y);
            unc = fin && (unc || (!failed && ue));
        }
    } else if constexpr (MODE == MODE_EVD) {
        const int band = (a.method == FRINGE_METHOD_STBAS) ? a.bw : -1;
        const bool fin = load_rows<T, Tg, NR>(h, grow, n, r, act, band);
        tridiag<T, NR, NC>(L, h, n, r);
        const T lam = bisect<T, NR>(L, n, n - 1, r, tn, pm);
        const T yr = tri_eigvec<T, NR>(L, n, r, lam, tn);
        y = back_transform<T, NR>(L, h, n, r, cx<T>{yr, T(0)});
        if (!fin) { code = -6.0f; failed = true; }
        else if (!(lam >= T(1.0e-6))) { code = -7.0f; failed = true; }
        if constexpr (CERT) {
            const bool ue = eig_uncertain<T, NR>(L, n, r, lam, tn, pm, true, y);
            unc = fin && (ue || !(lam >= T(1.0e-6) + T(kCertThrMargin) * Lim<T>::eps() * tn));
        }
    } else {
        double M[Blk<NR>::BR][8], m[NR];
        double pmin;
        bool fin = abs_gamma_block<NR, Tg>(M, gam, n, r, act, 0.0);
        fin = minR<NR>(fin ? 1.0f : 0.0f) > 0.5f;
        const bool pd = sweep_block<T, NR>(L, M, n, r, pmin);
        blocks_to_rows<T, NR>(L, M, m, r);
        const bool hfin = hadamard_rows<T, Tg, NR>(h, m, grow, n, r, act);
        const bool run_evd = !pd || !hfin;
        if constexpr (CERT) unc = !(pmin >= kCertPivotMin);
        // the EMI eigenpair, unless every pixel of the wave falls back to EVD
        // (a pixel's result never depends on whether this ran for it)
        if (__any(!run_evd)) {
            tridiag<T, NR, NC>(L, h, n, r);
            const T lam = bisect<T, NR>(L, n, 0, r, tn, pm);
            const T yr = tri_eigvec<T, NR>(L, n, r, lam, tn);
            y = back_transform<T, NR>(L, h, n, r, cx<T>{yr, T(0)});
            if constexpr (CERT) {
                // every lane runs the check (it holds cross-lane reductions)
                const bool ue = eig_uncertain<T, NR>(L, n, r, lam, tn, pm, false, y);
                unc = unc || (!run_evd && ue);
            }
        }
        if (__any(act && fin && run_evd)) {
            load_rows<T, Tg, NR>(h, grow, n, r, act, -1);
            tridiag<T, NR, NC>(L, h, n, r);
            const T le = bisect<T, NR>(L, n, n - 1, r, tn, pm);
            const T ye = tri_eigvec<T, NR>(L, n, r, le, tn);
            const cx<T> yv = back_transform<T, NR>(L, h, n, r, cx<T>{ye, T(0)});
            if constexpr (CERT) {
                const bool ue = eig_uncertain<T, NR>(L, n, r, le, tn, pm, true, yv);
                if (run_evd) unc = unc || ue;
            }
            if (run_evd) y = yv;
        }
        if (!fin) { code = -1.0f; failed = true; }
        if constexpr (CERT) unc = unc && fin;
    }

    // ---- referenced phases (evd.cpp:738-749) ----
    if (r == a.ref) L.YREF = y;
    wsync();
    const cx<T> vref = L.YREF;
    bool okr, oku, okv;
    const cx<T> cref = unit_or(cconj(vref), cx<T>{T(1), T(0)}, okr);
    const cx<T> u = unit_or(cmulc(y, vref), cx<T>{T(0), T(0)}, oku);
    cx<T> o = oku ? u : (okr ? cref : unit_or(y, cx<T>{T(1), T(0)}, okv));
    if (r == a.ref) o = cx<T>{T(1), T(0)};
    const cx<float> of = {(float)o.re, (float)o.im};
    L.OUT[r] = of;
    wsync();

    const bool write_ok = act && !failed;
    // ---- compressed SLC (evd.cpp:755-762) ----
    cx<double> cs = {0.0, 0.0};
    if (r >= a.ref && r < n && write_ok) {
        const float2 z = a.slc[(long long)r * a.npx + pix];
        cs.re = (double)(z.x * of.re + z.y * of.im);
        cs.im = (double)(z.y * of.re - z.x * of.im);
    }
    cs = sumR<NR>(cs);
    // ---- temporal coherence (evd.cpp:770-786) ----
    // sum over pairs i < j (j <= i + bw for STBAS) of
    //   exp(i(arg Gamma_ij - arg out_i + arg out_j)).
    // Pairs are dealt out circulantly so every lane takes about (n-1)/2 of
    // them: lane i forms t = g_ij conj(o_i) o_j with j = i + d (mod n) from
    // its own row of Gamma, and the pair's term is t for i < j, conj(t) for
    // i > j (Gamma and the phasor products are Hermitian).
    const bool stbas = (a.method == FRINGE_METHOD_STBAS);
    cx<T> tc = {T(0), T(0)};
    if (r < n && write_ok) {
        bool okx;
        const cx<T> ci = unit_or(cx<T>{(T)of.re, -(T)of.im}, cx<T>{T(1), T(0)}, okx);
        const int half = (n - 1) / 2;
        const int dmax = ((n & 1) == 0 && r < n / 2) ? half + 1 : half;
        // all gathers first, then the sums; jd[d-1] = (r + d) mod n, stepped
        // so that it stays in [0, n) for every n >= 1
        int jd[NR / 2];
        cx<Tg> gv[NR / 2];
#pragma unroll
        for (int d = 1; d <= NR / 2; ++d) {
            const int prev = d == 1 ? r : jd[d - 2];
            jd[d - 1] = prev + 1 == n ? 0 : prev + 1;
            gv[d - 1] = grow[jd[d - 1]];
        }
#pragma unroll
        for (int d = 1; d <= NR / 2; ++d) {
            const int j = jd[d - 1];
            const int lo = min(r, j), hi = max(r, j);
            const bool use = d <= dmax && !(stbas && hi - lo > a.bw);
            const cx<T> ph = unit_or(cx<T>{(T)gv[d - 1].re, (T)gv[d - 1].im}, cx<T>{T(1), T(0)}, okx);
            const cx<float> oj = L.OUT[j];
            cx<T> t = cmul(cmul(ph, ci), cx<T>{(T)oj.re, (T)oj.im});
            if (r > j) t = cconj(t);
            if (use) tc = cadd(tc, t);
        }
    }
    tc = sumR<NR>(tc);

    if (!inslot) return;
    if (r < n) a.phase[(long long)r * a.npx + pix] = write_ok ? make_float2(of.re, of.im) : make_float2(0.0f, 0.0f);
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
        if constexpr (CERT) {
            if (act && unc) {
                const int e = atomicAdd(a.cert_count, 1);
                if (e < a.cert_cap) a.cert_list[e] = (int)pix;
            }
        }
    }
}

// LIST = false: slot s of the round's coherence buffer is pixel first + s;
// NR = 32 holds two pixels per wave, NR = 64 one.  The FP32 instances
// certify their results.  LIST = true (T = Tg = double): the FP64 re-solve
// of the listed pixels; each group computes its pixel's coherence with the
// FP64 Gram into LDS and solves it there, looping over the list.  NC: the
// columns the tridiagonal reductions process (n <= NC; same results).
template <typename T, typename Tg, int MODE, int NR, bool LIST, int NC = NR>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && NR == 32 && MODE == MODE_MLE ? 4 : 1)))
void solve_reg_kernel(SolveArgs a) {
    __shared__ LdsR<T, NR> sm[64 / NR];
    const int grp = threadIdx.x / NR, r = threadIdx.x % NR;
    LdsR<T, NR>& L = sm[grp];
    if constexpr (!LIST) {
        // slots last-first: the Gram of this round has just written them in
        // tile order, and the most recently written coherence matrices are the
        // ones still in the last-level cache (measured 1.3 % on the solve)
        const int slot = (gridDim.x - 1 - blockIdx.x) * (64 / NR) + grp;
        const bool inslot = slot < a.slots;
        const int np = inslot ? a.npix[slot] : -1;
        const cx<Tg>* gam = (const cx<Tg>*)a.gamma + (size_t)(inslot ? slot : 0) * a.n * a.n;
        const long long pix = a.first + (inslot ? slot : 0);
        solve_reg_pixel<T, Tg, MODE, NR, sizeof(T) == 4, NC>(a, L, r, inslot, np, gam, pix);
    } else {
        static_assert(sizeof(T) == 8 && sizeof(Tg) == 8, "the re-solve runs in FP64");
        constexpr int GPW = 64 / NR;
        extern __shared__ __align__(16) char dyn[];
        // the group's coherence in its global area (a.awork), the Gram's
        // scratch in LDS
        cx<double>* gl = (cx<double>*)((char*)a.awork + a.awork_stride * ((size_t)blockIdx.x * GPW + grp));
        char* gscratch = dyn + grp * a.group_bytes;
        const int cnt = min(*a.list_count, a.list_cap);
        for (int b0 = blockIdx.x * GPW; b0 < cnt; b0 += gridDim.x * GPW) {
            const int e = b0 + grp;
            const bool valid = e < cnt;
            const long long pix = valid ? (long long)a.list[e] : 0;
            const int np = gram_f64_pixel<NR, 1, kResolveSmallCB>(a.g64, valid, pix, gl, gscratch);
            __threadfence();
            __syncthreads();
            solve_reg_pixel<T, Tg, MODE, NR, false, NC>(a, L, r, valid, np, (const cx<Tg>*)gl, pix);
            __syncthreads();
        }
    }
}

// Persistent launch up to 64 dates: each wave takes units (two pixels for
// NR = 32, one for NR = 64) of the queued jobs until the queue is drained.
template <typename T, typename Tg, int MODE, int NR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && NR == 32 ? 4 : 1)))
void solve_reg_persistent(PersistArgs pa) {
    __shared__ LdsR<T, NR> sm[64 / NR];
    __shared__ int s_u;
    const int total = pa.unit_end[pa.njobs - 1];
    for (;;) {
        const int u0 = take_units(pa, total, &s_u);
        if (u0 >= total) break;
        const int u1 = min(u0 + pa.grab, total);
        for (int v = u0; v < u1; ++v) {
            // the lane's row, opaque to the compiler on every pixel: values derived
            // from it are then computed inside the loop instead of being hoisted
            // out of it and held in registers across the whole estimator chain
            int r = threadIdx.x % NR, grp = threadIdx.x / NR;
            asm volatile("" : "+v"(r), "+v"(grp));
            LdsR<T, NR>& L = sm[grp];
            const int u = total - 1 - v;
            const int j = job_of_unit(pa, u);
            const PJob& J = pa.job[j];
            const SolveArgs a = job_args(pa, J);
            const int slot = (u - (j ? pa.unit_end[j - 1] : 0)) * (64 / NR) + grp;
            const bool inslot = slot < a.slots;
            const int np = inslot ? a.npix[slot] : -1;
            const cx<Tg>* gam = (const cx<Tg>*)a.gamma + (size_t)(inslot ? slot : 0) * a.n * a.n;
            solve_reg_pixel<T, Tg, MODE, NR, sizeof(T) == 4>(a, L, r, inslot, np, gam, a.first + (inslot ? slot : 0));
        }
    }
}

}  // namespace fringe
