// link_solve.hpp -- the per-pixel estimator chain that follows the Gram:
// normalised coherence Gamma -> (|Gamma|^-1 o Gamma) -> extreme eigenpair ->
// referenced phases, compressed SLC, temporal coherence.
//
// Lane groups: a wave holds 64/G pixels (G <= 64), or a workgroup of G/64
// waves holds one (G > 64); lane g of a group owns matrix rows g + G*q
// (q < R), so n <= G*R.  All control flow that contains a barrier is uniform
// over the workgroup (it depends only on launch arguments or on group-wide
// votes); per-pixel decisions are carried as predicates.
#pragma once
#include "device_common.hpp"
#include "fringe_amd.h"
#include "gram64.hpp"

namespace fringe {

struct SolveArgs {
    const void* gamma;   // [slots][n][n] cx<Tg>, row-major, Hermitian
    const int* npix;     // [slots]: neighbour count, -1 = pixel left untouched
    const float2* slc;   // tile SLC stack [band][row][col] (centre pixel for comp)
    long long npx;       // rows*cols of the tile
    long long first;     // linear pixel index of slot 0
    int slots;
    int n, module, method, ref, bw;
    size_t group_bytes;
    float2* phase;       // [band][row][col]
    float* tcorr;        // [row][col]
    float2* comp;        // [row][col]
    // FP32 solver: linear pixel indices whose result it cannot certify to the
    // phase tolerance are appended here (cert_count entries, at most cert_cap)
    int* cert_list;
    int* cert_count;
    int cert_cap;
    // FP64 re-solve of a pixel list: entry e is pixel list[e], e < *list_count
    // (at most list_cap); the coherence is recomputed by the FP64 Gram (g64)
    const int* list;
    const int* list_count;
    int list_cap;
    GramArgs g64;
    // Above 64 dates, when a pixel's matrices do not fit a workgroup's LDS:
    // the solver's n x (n|1) work matrix lives in global memory, awork +
    // awork_stride * (slot, or workgroup in the re-solve); the re-solve also
    // keeps its FP64 coherence there, at the start of each workgroup's area.
    void* awork;
    size_t awork_stride;
};

// Certification of an FP32 eigenpair.  The FP32 chain's backward error is
// C eps32 |H| for a small C, so the selected unit eigenvector v moves by
// about C eps32 |H| / gap, and the phase of its smallest component by that
// over |v_t|.  Against the FP64 oracle, pixels with small gaps and small
// components showed C between 0.3 and 2.4 (|H| taken as the spectral spread;
// |T| below bounds it from above).  A result counts as certified when the
// estimate with C = 5 stays below 5e-4 rad, half the phase tolerance: no
// eigenvalue of T may lie within 1e4 eps32 |T| / min|v_t| of the selected
// one.  Decisions against the reference's 1e-6 eigenvalue thresholds are
// certified 256 eps32 |T| away from them, and the positive definiteness of
// |Gamma| when no sweep pivot up to the first non-positive one lies within
// 1e-4 of zero (|Gamma| has unit diagonal, so its pivots lie in (0, 1] when
// it is definite).
constexpr float kCertScale = 1.0e4f;
constexpr float kCertThrMargin = 256.0f;
constexpr double kCertPivotMin = 1.0e-4;

// ---- persistent solve launch over a queue of jobs -------------------------
// A job is one round segment of one tile: slots [0, slots) of its coherence
// buffer are pixels first + s of that tile.  One launch drains the work
// units of up to kMaxJobs jobs, of different date counts and tiles, from an
// atomic counter (a unit is one wave's share: two slots for n <= 32, one
// slot otherwise); units are taken last-first, the matrices the Gram wrote
// last being the ones still cached.
constexpr int kMaxJobs = 16;
struct PJob {
    const void* gamma;
    const int* npix;
    const float2* slc;
    float2* phase;
    float* tcorr;
    float2* comp;
    int* cert_list;
    int* cert_count;
    void* awork;        // above 64 dates: work matrices in global memory (or null)
    size_t awork_stride;
    long long npx, first;
    int slots, n, ref, bw, cert_cap;
    unsigned group_bytes;   // above 64 dates: dynamic LDS of this job's group
};
struct PersistArgs {
    int module, method, njobs, grab;
    int* queue;                 // 0 at launch
    int unit_end[kMaxJobs];     // prefix sums of the jobs' units
    PJob job[kMaxJobs];
};

template <typename T> struct Lim;
template <> struct Lim<float> {
    static __device__ constexpr float eps() { return 5.9604645e-8f; }
    static __device__ constexpr float tiny() { return 1.0e-30f; }
    static __device__ constexpr int iters(int G) { return G >= 64 ? 6 : 7; }
};
template <> struct Lim<double> {
    static __device__ constexpr double eps() { return 1.1102230246251565e-16; }
    static __device__ constexpr double tiny() { return 1.0e-290; }
    static __device__ constexpr int iters(int G) { return G >= 64 ? 10 : 12; }
};

// Per-group LDS carve-up.
template <typename T> struct SolveLds {
    cx<T>* A;        // n x S (S = n|1), also hosts the FP64 |Gamma| inverse
    cx<T>* V;        // reflector being applied
    cx<T>* W;
    cx<T>* TAU;      // reflector scalars
    cx<T>* Y;        // eigenvector, for the reference lookup
    T *D, *E, *E2;   // tridiagonal T
    T *X, *DD, *DU, *DU2, *DL;  // inverse-iteration work
    cx<float>* OUT;  // referenced unit phasors
};

__host__ __device__ inline size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }

// LDS bytes of one pixel's group; a_lds = false leaves the work matrix A
// out (it is then in global memory).
template <typename T> __host__ __device__ inline size_t solve_group_bytes(int n, bool a_lds = true) {
    const int S = n | 1;
    size_t b = a_lds ? align16((size_t)n * S * sizeof(cx<T>)) : 0;
    b += 4 * align16((size_t)n * sizeof(cx<T>));
    b += 8 * align16((size_t)n * sizeof(T));
    b += align16((size_t)n * sizeof(cx<float>));
    return b;
}

template <typename T> __device__ inline SolveLds<T> solve_carve(char* base, int n, cx<T>* aglob = nullptr) {
    SolveLds<T> L;
    const int S = n | 1;
    char* p = base;
    if (aglob) {
        L.A = aglob;
    } else {
        L.A = (cx<T>*)p;
        p += align16((size_t)n * S * sizeof(cx<T>));
    }
    L.V = (cx<T>*)p;   p += align16((size_t)n * sizeof(cx<T>));
    L.W = (cx<T>*)p;   p += align16((size_t)n * sizeof(cx<T>));
    L.TAU = (cx<T>*)p; p += align16((size_t)n * sizeof(cx<T>));
    L.Y = (cx<T>*)p;   p += align16((size_t)n * sizeof(cx<T>));
    L.D = (T*)p;   p += align16((size_t)n * sizeof(T));
    L.E = (T*)p;   p += align16((size_t)n * sizeof(T));
    L.E2 = (T*)p;  p += align16((size_t)n * sizeof(T));
    L.X = (T*)p;   p += align16((size_t)n * sizeof(T));
    L.DD = (T*)p;  p += align16((size_t)n * sizeof(T));
    L.DU = (T*)p;  p += align16((size_t)n * sizeof(T));
    L.DU2 = (T*)p; p += align16((size_t)n * sizeof(T));
    L.DL = (T*)p;  p += align16((size_t)n * sizeof(T));
    L.OUT = (cx<float>*)p;
    return L;
}

template <int G> __device__ inline bool group_all(bool b) {
    if constexpr (G > 64) return gmin<G>(b ? 1.0f : 0.0f) > 0.5f;   // through the shared scratch
    else return group_ballot<G>(!b) == 0ull;
}

// ---- load Gamma (optionally banded for STBAS) into A --------------------
template <typename T, typename Tg, int G, int R>
__device__ bool load_gamma(cx<T>* A, int S, const cx<Tg>* gam, int n, bool act, int band) {
    const int g = threadIdx.x % G;
    bool fin = true;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int r = g + G * q;
        if (r >= n) continue;
        for (int c = 0; c < n; ++c) {
            cx<Tg> v = act ? gam[(size_t)r * n + c] : cx<Tg>{Tg(r == c ? 1 : 0), Tg(0)};
            if (band >= 0 && (c - r > band || r - c > band)) v = cx<Tg>{Tg(0), Tg(0)};
            fin = fin && isfinite(v.re) && isfinite(v.im);
            A[r * S + c] = cx<T>{(T)v.re, (T)v.im};
        }
    }
    __syncthreads();
    return group_all<G>(fin);
}

// ---- |Gamma| (unit diagonal, minus shift) as FP64 in A's storage ---------
// M row r starts where A row r starts; MS is the row stride in doubles.
template <typename Tg, int G, int R>
__device__ void load_abs_gamma(double* M, int MS, const cx<Tg>* gam, int n, bool act, double shift) {
    const int g = threadIdx.x % G;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int r = g + G * q;
        if (r >= n) continue;
        for (int c = 0; c < n; ++c) {
            double m;
            if (c == r) {
                m = 1.0 - shift;
            } else if (act) {
                const cx<Tg> v = gam[(size_t)r * n + c];
                m = hypot((double)v.re, (double)v.im);
            } else {
                m = 0.0;
            }
            M[r * MS + c] = m;
        }
    }
    __syncthreads();
}

// ---- in-place inverse of a symmetric positive definite matrix ------------
// Gauss-Jordan sweep without pivoting.  The pivot met at step j is the
// Schur complement M_jj - M_j,<j M_<j,<j^-1 M_<j,j, i.e. the square of the
// j-th Cholesky diagonal, so the matrix is declared not positive definite
// exactly when a Cholesky factorisation would be (a pivot <= 0 or NaN).
template <int G, int R>
__device__ bool sweep_inverse(double* M, int MS, int n, double& pmin) {
    const int g = threadIdx.x % G;
    bool ok = true;
    pmin = 1.0;
    for (int j = 0; j < n; ++j) {
        const double p = M[j * MS + j];
        if (ok) pmin = fmin(pmin, fabs(p));   // up to the first pivot that is not positive
        ok = ok && (p > 0.0);
        const double piv = 1.0 / p;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int c = g + G * q;
            if (c < n) M[j * MS + c] = (c == j) ? piv : M[j * MS + c] * piv;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = g + G * q;
            if (r >= n || r == j) continue;
            const double f = M[r * MS + j];
            for (int c = 0; c < n; ++c)
                if (c != j) M[r * MS + c] -= f * M[j * MS + c];
            M[r * MS + j] = -f * piv;
        }
        __syncthreads();
    }
    // mirror the upper triangle into the lower one (zpotri leaves only the
    // upper triangle and EigenLapack.hpp:240-248 mirrors it the same way)
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int r = g + G * q;
        if (r >= n) continue;
        for (int c = 0; c < r; ++c) M[r * MS + c] = M[c * MS + r];
    }
    __syncthreads();
    return group_all<G>(ok);
}

// ---- H = M o Gamma into A (decreasing columns keep FP64 M intact until read)
template <typename T, typename Tg, int G, int R>
__device__ bool form_hadamard(cx<T>* A, int S, const double* M, int MS, const cx<Tg>* gam, int n,
                              bool act) {
    const int g = threadIdx.x % G;
    bool fin = true;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int r = g + G * q;
        if (r >= n) continue;
        for (int c = n - 1; c >= 0; --c) {
            const double m = M[r * MS + c];
            const cx<Tg> v = act ? gam[(size_t)r * n + c] : cx<Tg>{Tg(r == c ? 1 : 0), Tg(0)};
            const double hr = m * (double)v.re, hi = m * (double)v.im;
            fin = fin && isfinite(hr) && isfinite(hi);
            A[r * S + c] = cx<T>{(T)hr, (T)hi};
        }
    }
    __syncthreads();
    return group_all<G>(fin);
}

// ---- Householder tridiagonalisation A = Q T Q^H ---------------------------
// Reflector j maps column j below the diagonal onto beta*e1 with beta real,
// so T = tridiag(D, E) is real.  v_j is stored in A[r][j] for r > j+1
// (v_{j+1} = 1 implicit) and tau_j in TAU[j].
template <typename T, int G, int R>
__device__ void tridiagonalize(SolveLds<T>& L, int S, int n) {
    const int g = threadIdx.x % G;
    cx<T>* A = L.A;
    for (int j = 0; j < n - 1; ++j) {
        const cx<T> alpha = A[(j + 1) * S + j];
        T xn2 = T(0);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = g + G * q;
            if (r > j + 1 && r < n) xn2 += cabs2(A[r * S + j]);
        }
        xn2 = gsum<G>(xn2);
        const T ar = alpha.re, ai = alpha.im;
        cx<T> tau = {T(0), T(0)}, scal = {T(0), T(0)};
        T beta = ar;
        if (xn2 != T(0) || ai != T(0)) {
            const T nrm = sqrt(ar * ar + ai * ai + xn2);
            beta = (ar >= T(0)) ? -nrm : nrm;
            tau = cx<T>{(beta - ar) / beta, -ai / beta};
            scal = crecip(cx<T>{ar - beta, ai});
        }
        if (g == 0) {
            L.D[j] = A[j * S + j].re;
            L.E[j] = beta;
            L.TAU[j] = tau;
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = g + G * q;
            if (r >= n) continue;
            cx<T> v = {T(0), T(0)};
            if (r == j + 1) {
                v = cx<T>{T(1), T(0)};
            } else if (r > j + 1) {
                v = cmul(A[r * S + j], scal);
                A[r * S + j] = v;
            }
            L.V[r] = v;
        }
        __syncthreads();
        // p = tau * A22 v ; w = p - (tau/2)(p^H v) v
        cx<T> p[R];
        cx<T> s = {T(0), T(0)};
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = g + G * q;
            p[q] = cx<T>{T(0), T(0)};
            if (r > j && r < n) {
                for (int c = j + 1; c < n; ++c) cfma(p[q], A[r * S + c], L.V[c]);
                p[q] = cmul(tau, p[q]);
                const cx<T> vr = L.V[r];
                s = cadd(s, cmulc(vr, p[q]));
            }
        }
        s = gsum<G>(s);
        const cx<T> alpha2 = cscale(cmul(tau, s), T(-0.5));
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = g + G * q;
            if (r > j && r < n) L.W[r] = cadd(p[q], cmul(alpha2, L.V[r]));
        }
        __syncthreads();
        // A22 -= v w^H + w v^H
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = g + G * q;
            if (r <= j || r >= n) continue;
            const cx<T> vr = L.V[r], wr = L.W[r];
            for (int c = j + 1; c < n; ++c) {
                cx<T> a = A[r * S + c];
                cfms_c(a, vr, L.W[c]);
                cfms_c(a, wr, L.V[c]);
                A[r * S + c] = a;
            }
        }
        __syncthreads();
    }
    if (g == 0) L.D[n - 1] = A[(n - 1) * S + (n - 1)].re;
    __syncthreads();
}

// Number of eigenvalues of tridiag(D, E) below sigma (LDL^T inertia).
template <typename T>
__device__ inline int sturm_count(const T* D, const T* E2, int n, T sigma, T pivmin) {
    T q = D[0] - sigma;
    if (fabs(q) < pivmin) q = -pivmin;
    int cnt = (q < T(0)) ? 1 : 0;
    for (int i = 1; i < n; ++i) {
        q = (D[i] - sigma) - E2[i - 1] / q;
        if (fabs(q) < pivmin) q = -pivmin;
        cnt += (q < T(0)) ? 1 : 0;
    }
    return cnt;
}

// k-th smallest eigenvalue (0-based) by (G+1)-section: lane g tests the
// g-th interior point of the current interval.  tnorm receives max|D|+|E|.
template <typename T, int G>
__device__ T select_eigenvalue(SolveLds<T>& L, int n, int k, T& tnorm, T& pm) {
    const int g = threadIdx.x % G;
    T lo = T(3.0e38), hi = T(-3.0e38), em2 = T(0), tn = T(0);
    for (int i = g; i < n; i += G) {
        const T a = (i > 0) ? fabs(L.E[i - 1]) : T(0);
        const T b = (i < n - 1) ? fabs(L.E[i]) : T(0);
        lo = min(lo, L.D[i] - a - b);
        hi = max(hi, L.D[i] + a + b);
        tn = max(tn, fabs(L.D[i]) + a + b);
        if (i < n - 1) {
            L.E2[i] = L.E[i] * L.E[i];
            em2 = max(em2, L.E2[i]);
        }
    }
    lo = gmin<G>(lo);
    hi = gmax<G>(hi);
    em2 = gmax<G>(em2);
    tn = gmax<G>(tn);
    __syncthreads();
    tnorm = tn;
    const T pivmin = Lim<T>::tiny() * max(T(1), em2);
    pm = pivmin;
    const T pad = T(4) * Lim<T>::eps() * tn + T(2) * pivmin;
    lo -= pad;
    hi += pad;
    const int iters = Lim<T>::iters(G);
    for (int it = 0; it < iters; ++it) {
        const T w = hi - lo;
        const T sig = lo + (w * T(g + 1)) / T(G + 1);
        const int cnt = sturm_count(L.D, L.E2, n, sig, pivmin);
        // first lane whose shift has more than k eigenvalues below it; the
        // counts grow with the shift, so that is G minus the number of such lanes
        int f;
        if constexpr (G > 64) {
            f = G - (int)gsum<G>(cnt > k ? 1.0f : 0.0f);
        } else {
            const uint64_t b = group_ballot<G>(cnt > k);
            f = b ? (int)__builtin_ctzll(b) : G;
        }
        const T nhi = (f < G) ? lo + (w * T(f + 1)) / T(G + 1) : hi;
        const T nlo = (f > 0) ? lo + (w * T(f)) / T(G + 1) : lo;
        lo = nlo;
        hi = nhi;
    }
    return T(0.5) * (lo + hi);
}

// Inverse iteration on tridiag(D, E) - lam I (Gaussian elimination with
// partial pivoting, small pivots lifted to eps*|T|); run by one lane.
template <typename T>
__device__ void tridiag_inverse_iteration(SolveLds<T>& L, int n, T lam, T tnorm) {
    T* X = L.X;
    T *DD = L.DD, *DU = L.DU, *DU2 = L.DU2, *DL = L.DL;
    const T pert = max(Lim<T>::eps() * tnorm, Lim<T>::tiny());
    for (int i = 0; i < n; ++i) X[i] = T(0.5) + T((i * 37 + 11) & 63) * T(1.0 / 64.0);
    for (int it = 0; it < 3; ++it) {
        for (int i = 0; i < n; ++i) {
            DD[i] = L.D[i] - lam;
            if (i < n - 1) {
                DU[i] = L.E[i];
                DL[i] = L.E[i];
            }
        }
        for (int i = 0; i < n - 1; ++i) {
            if (fabs(DD[i]) >= fabs(DL[i])) {
                if (fabs(DD[i]) < pert) DD[i] = (DD[i] < T(0)) ? -pert : pert;
                const T f = DL[i] / DD[i];
                DD[i + 1] -= f * DU[i];
                X[i + 1] -= f * X[i];
                DU2[i] = T(0);
            } else {
                const T f = DD[i] / DL[i];
                DD[i] = DL[i];
                const T tmp = DD[i + 1];
                DD[i + 1] = DU[i] - f * tmp;
                DU[i] = tmp;
                if (i < n - 2) {
                    DU2[i] = DU[i + 1];
                    DU[i + 1] = -f * DU2[i];
                } else {
                    DU2[i] = T(0);
                }
                const T tx = X[i];
                X[i] = X[i + 1];
                X[i + 1] = tx - f * X[i];
            }
        }
        if (fabs(DD[n - 1]) < pert) DD[n - 1] = (DD[n - 1] < T(0)) ? -pert : pert;
        X[n - 1] /= DD[n - 1];
        if (n > 1) X[n - 2] = (X[n - 2] - DU[n - 2] * X[n - 1]) / DD[n - 2];
        for (int i = n - 3; i >= 0; --i)
            X[i] = (X[i] - DU[i] * X[i + 1] - DU2[i] * X[i + 2]) / DD[i];
        T mx = T(0);
        for (int i = 0; i < n; ++i) mx = max(mx, fabs(X[i]));
        const T sc = (mx > T(0) && isfinite(mx)) ? T(1) / mx : T(1);
        for (int i = 0; i < n; ++i) X[i] *= sc;
    }
}

// Eigenvector of A for eigenvalue index k (0-based ascending) after
// tridiagonalize(): returns the eigenvalue, y holds the lane's rows.
template <typename T, int G, int R>
__device__ T extreme_eigenpair(SolveLds<T>& L, int S, int n, int k, cx<T> (&y)[R], T& tnorm, T& pivmin) {
    const int g = threadIdx.x % G;
    const T lam = select_eigenvalue<T, G>(L, n, k, tnorm, pivmin);
    if (g == 0) tridiag_inverse_iteration<T>(L, n, lam, tnorm);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int r = g + G * q;
        y[q] = cx<T>{(r < n) ? L.X[r] : T(0), T(0)};
    }
    // y <- Q y = H_0 H_1 ... H_{n-2} y
    for (int j = n - 2; j >= 0; --j) {
        cx<T> s = {T(0), T(0)};
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = g + G * q;
            if (r >= j + 1 && r < n) {
                const cx<T> v = (r == j + 1) ? cx<T>{T(1), T(0)} : L.A[r * S + j];
                s = cadd(s, cmulc(y[q], v));
            }
        }
        s = gsum<G>(s);
        const cx<T> ts = cmul(L.TAU[j], s);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = g + G * q;
            if (r >= j + 1 && r < n) {
                const cx<T> v = (r == j + 1) ? cx<T>{T(1), T(0)} : L.A[r * S + j];
                y[q] = csub(y[q], cmul(v, ts));
            }
        }
    }
    __syncthreads();
    return lam;
}

template <typename T> __device__ inline cx<T> unit_or(cx<T> u, cx<T> fallback, bool& ok) {
    const T a = sqrt(cabs2(u));
    ok = a > T(0);
    return ok ? cscale(u, T(1) / a) : fallback;
}

// Certification of the eigenpair (criteria at kCertScale): true when the
// unit eigenvector y (lane rows) of eigenvalue lam of tridiag(D, E), norm
// bound tn, has another eigenvalue within kCertScale eps |T| / min|v_t|, or
// a zero or non-finite component.  Every lane of the group takes part.
template <typename T, int G, int R>
__device__ bool eig_uncertain_g(const SolveLds<T>& L, int n, T lam, T tn, T pivmin, bool largest,
                                const cx<T> (&y)[R]) {
    const int g = threadIdx.x % G;
    T s2 = T(0), m2 = T(3.0e38);
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int r = g + G * q;
        if (r < n) {
            const T a2 = cabs2(y[q]);
            s2 += a2;
            m2 = min(m2, a2);
        }
    }
    s2 = gsum<G>(s2);
    m2 = gmin<G>(m2);
    const T vmin = sqrt(m2 / s2);
    const T delta = T(kCertScale) * Lim<T>::eps() * tn / vmin;
    if (!(delta <= tn)) return true;
    const int c = sturm_count(L.D, L.E2, n, largest ? lam - delta : lam + delta, pivmin);
    return largest ? (c <= n - 2) : (c >= 2);
}

// One pixel's estimator chain on the G lanes of its group (n > 64 in this
// build): gam is its coherence, np its neighbour count (-1: not linked).
// CERT: the FP32 result is checked and, when it is not certified, the pixel
// is appended to a.cert_list for the FP64 re-solve.
template <typename T, typename Tg, int G, int R, bool CERT>
__device__ __forceinline__ void solve_generic_pixel(const SolveArgs& a, SolveLds<T>& L, bool inslot, int np,
                                                    const cx<Tg>* gam, long long pix) {
    const int g = threadIdx.x % G;
    const int n = a.n, S = n | 1;
    const int MS = S * (int)(sizeof(cx<T>) / sizeof(double));
    double* M = (double*)L.A;
    const bool act = np >= 0;

    float code = 0.0f;
    bool failed = false;
    cx<T> y[R];
    T tn, pm;
    bool unc = false;   // CERT: the FP32 result is not certified

    if (a.module == FRINGE_MODULE_EVD && a.method == FRINGE_METHOD_MLE) {
        // evd.cpp:584-688: lambda_min(Gamma) (-1/-2), lambda_min(|Gamma|) (-3/-4),
        // inv(|Gamma|) (-5), smallest eigenpair of inv(|Gamma|) o Gamma (-6/-7).
        const bool fin = load_gamma<T, Tg, G, R>(L.A, S, gam, n, act, -1);
        tridiagonalize<T, G, R>(L, S, n);
        const T lmin = select_eigenvalue<T, G>(L, n, 0, tn, pm);
        if (!fin) { code = -1.0f; failed = true; }
        else if (!(lmin >= T(1.0e-6))) { code = -2.0f; failed = true; }
        if constexpr (CERT) unc = !(lmin >= T(1.0e-6) + T(kCertThrMargin) * Lim<T>::eps() * tn);
        double pmin;
        load_abs_gamma<Tg, G, R>(M, MS, gam, n, act, 1.0e-6);
        const bool pd_shift = sweep_inverse<G, R>(M, MS, n, pmin);
        if (!failed && !pd_shift) { code = -4.0f; failed = true; }
        if constexpr (CERT) unc = unc || !(pmin >= kCertPivotMin);
        load_abs_gamma<Tg, G, R>(M, MS, gam, n, act, 0.0);
        const bool pd = sweep_inverse<G, R>(M, MS, n, pmin);
        if (!failed && !pd) { code = -5.0f; failed = true; }
        if constexpr (CERT) unc = unc || !(pmin >= kCertPivotMin);
        const bool hfin = form_hadamard<T, Tg, G, R>(L.A, S, M, MS, gam, n, act);
        tridiagonalize<T, G, R>(L, S, n);
        const T lam = extreme_eigenpair<T, G, R>(L, S, n, 0, y, tn, pm);
        if (!failed && !hfin) { code = -6.0f; failed = true; }
        if (!failed && !(lam >= T(1.0e-6))) { code = -7.0f; failed = true; }
        if constexpr (CERT) {
            unc = unc || !(lam >= T(1.0e-6) + T(kCertThrMargin) * Lim<T>::eps() * tn);
            const bool ue = eig_uncertain_g<T, G, R>(L, n, lam, tn, pm, false, y);
            unc = fin && (unc || (!failed && ue));
        }
    } else if (a.module == FRINGE_MODULE_EVD) {
        // EVD / STBAS (evd.cpp:689-732)
        const int band = (a.method == FRINGE_METHOD_STBAS) ? a.bw : -1;
        const bool fin = load_gamma<T, Tg, G, R>(L.A, S, gam, n, act, band);
        tridiagonalize<T, G, R>(L, S, n);
        const T lam = extreme_eigenpair<T, G, R>(L, S, n, n - 1, y, tn, pm);
        if (!fin) { code = -6.0f; failed = true; }
        else if (!(lam >= T(1.0e-6))) { code = -7.0f; failed = true; }
        if constexpr (CERT) {
            const bool ue = eig_uncertain_g<T, G, R>(L, n, lam, tn, pm, true, y);
            unc = fin && (ue || !(lam >= T(1.0e-6) + T(kCertThrMargin) * Lim<T>::eps() * tn));
        }
    } else {
        // phase_link EMI (phase_link.cpp:540-618)
        bool fin = true;
        load_abs_gamma<Tg, G, R>(M, MS, gam, n, act, 0.0);
        {
            // Gamma must be finite for the reference's first zheevr to succeed (-1)
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int r = g + G * q;
                if (r >= n) continue;
                for (int c = 0; c < n; ++c) fin = fin && isfinite(M[r * MS + c]);
            }
            fin = group_all<G>(fin);
        }
        double pmin;
        const bool pd = sweep_inverse<G, R>(M, MS, n, pmin);
        const bool hfin = form_hadamard<T, Tg, G, R>(L.A, S, M, MS, gam, n, act);
        // run_evd is uniform over the group: the barriers below are block-wide
        bool run_evd = !pd || !hfin;
        if constexpr (CERT) unc = !(pmin >= kCertPivotMin);
        // the EMI eigenpair, unless the pixel falls back to EVD (the test is
        // over the whole workgroup, so the barriers inside stay uniform)
        if (__syncthreads_or(!run_evd)) {
            tridiagonalize<T, G, R>(L, S, n);
            const T lam = extreme_eigenpair<T, G, R>(L, S, n, 0, y, tn, pm);
            if constexpr (CERT) {
                const bool ue = eig_uncertain_g<T, G, R>(L, n, lam, tn, pm, false, y);
                unc = unc || (!run_evd && ue);
            }
        }
        if (__any(act && fin && run_evd)) {
            // EVD fallback: largest eigenvector of Gamma
            load_gamma<T, Tg, G, R>(L.A, S, gam, n, act, -1);
            tridiagonalize<T, G, R>(L, S, n);
            cx<T> ye[R];
            const T le = extreme_eigenpair<T, G, R>(L, S, n, n - 1, ye, tn, pm);
            if constexpr (CERT) {
                const bool ue = eig_uncertain_g<T, G, R>(L, n, le, tn, pm, true, ye);
                if (run_evd) unc = unc || ue;
            }
            if (run_evd) {
#pragma unroll
                for (int q = 0; q < R; ++q) y[q] = ye[q];
            }
        }
        if (!fin) { code = -1.0f; failed = true; }
        if constexpr (CERT) unc = unc && fin;
    }

    // ---- referenced phases (evd.cpp:738-749) ------------------------------
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int r = g + G * q;
        if (r < n) L.Y[r] = y[q];
    }
    __syncthreads();
    const cx<T> vref = L.Y[a.ref];
    bool okr;
    const cx<T> cref = unit_or(cconj(vref), cx<T>{T(1), T(0)}, okr);
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int r = g + G * q;
        if (r >= n) continue;
        bool oku, okv;
        const cx<T> u = unit_or(cmulc(y[q], vref), cx<T>{T(0), T(0)}, oku);
        cx<T> o;
        if (oku) o = u;
        else if (okr) o = cref;
        else o = unit_or(y[q], cx<T>{T(1), T(0)}, okv);
        if (r == a.ref) o = cx<T>{T(1), T(0)};
        L.OUT[r] = cx<float>{(float)o.re, (float)o.im};
    }
    __syncthreads();

    const bool write_ok = act && !failed;
    // ---- compressed SLC from the centre pixel (evd.cpp:755-762) ----------
    cx<double> cs = {0.0, 0.0};
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int r = g + G * q;
        if (r >= a.ref && r < n && write_ok) {
            const float2 z = a.slc[(long long)r * a.npx + pix];
            const cx<float> o = L.OUT[r];
            const float pr = z.x * o.re + z.y * o.im, pi = z.y * o.re - z.x * o.im;
            cs.re += (double)pr;
            cs.im += (double)pi;
        }
    }
    cs = gsum<G>(cs);
    // ---- temporal coherence (evd.cpp:770-786) -----------------------------
    const bool stbas = (a.method == FRINGE_METHOD_STBAS);
    cx<double> tc = {0.0, 0.0};
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int i = g + G * q;
        if (i >= n || !write_ok) continue;
        const int ulim = stbas ? min(n, i + a.bw + 1) : n;
        bool okx;
        const cx<float> oi = L.OUT[i];
        const cx<double> ci = unit_or(cx<double>{(double)oi.re, -(double)oi.im}, cx<double>{1.0, 0.0}, okx);
        for (int j = i + 1; j < ulim; ++j) {
            const cx<Tg> gv = gam[(size_t)i * n + j];
            const cx<double> ph = unit_or(cx<double>{(double)gv.re, (double)gv.im}, cx<double>{1.0, 0.0}, okx);
            const cx<float> oj = L.OUT[j];
            const cx<double> cj = unit_or(cx<double>{(double)oj.re, (double)oj.im}, cx<double>{1.0, 0.0}, okx);
            tc = cadd(tc, cmul(cmul(ph, ci), cj));
        }
    }
    tc = gsum<G>(tc);

    if (!inslot) return;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int r = g + G * q;
        if (r >= n) continue;
        const cx<float> o = write_ok ? L.OUT[r] : cx<float>{0.0f, 0.0f};
        a.phase[(long long)r * a.npx + pix] = make_float2(o.re, o.im);
    }
    if (g == 0) {
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
            tco = (float)(sqrt(tc.re * tc.re + tc.im * tc.im) / (double)counter);
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

// Solver arguments of job J for the shared per-pixel bodies.
__device__ inline SolveArgs job_args(const PersistArgs& pa, const PJob& J) {
    SolveArgs a{};
    a.gamma = J.gamma;
    a.npix = J.npix;
    a.slc = J.slc;
    a.npx = J.npx;
    a.first = J.first;
    a.slots = J.slots;
    a.n = J.n;
    a.module = pa.module;
    a.method = pa.method;
    a.ref = J.ref;
    a.bw = J.bw;
    a.group_bytes = J.group_bytes;
    a.phase = J.phase;
    a.tcorr = J.tcorr;
    a.comp = J.comp;
    a.cert_list = J.cert_list;
    a.cert_count = J.cert_count;
    a.cert_cap = J.cert_cap;
    a.awork = J.awork;
    a.awork_stride = J.awork_stride;
    return a;
}

// Next unit range of this workgroup from the queue: [u0, u1) counted from
// the end; u0 >= total when the queue is drained.
__device__ inline int take_units(const PersistArgs& pa, int total, int* s_u) {
    __syncthreads();
    if (threadIdx.x == 0) *s_u = atomicAdd(pa.queue, pa.grab);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*s_u);   // uniform: keeps the job's fields in scalar registers
}

__device__ inline int job_of_unit(const PersistArgs& pa, int u) {
    int j = 0;
    while (j + 1 < pa.njobs && u >= pa.unit_end[j]) ++j;
    return __builtin_amdgcn_readfirstlane(j);
}

// Above 64 dates.  LIST = false: one pixel per group, slot s of the round's
// coherence buffer is pixel first + s; the FP32 instance certifies its
// results.  LIST = true (T = Tg = double, G = 128): the FP64 re-solve of the
// listed pixels; both waves of a workgroup compute the listed pixel's FP64
// coherence (the same values) into the workgroup's global area, then solve
// it, looping over the list.
// Column block of the re-solve's FP64 Gram, the plain path's 16.  (Each
// coherence entry sums the same products in the same order whatever the
// block; smaller blocks, which let the re-solve run three waves per SIMD
// instead of one, cost more in passes, two barriers per position each,
// than the occupancy gained: measured slower.)
constexpr int kResolveCB = 16;

template <typename T, typename Tg, int G, int R, bool LIST>
__global__ __launch_bounds__(G > 64 ? G : 64) void solve_kernel(SolveArgs a) {
    constexpr int GPW = G >= 64 ? 1 : 64 / G;
    const int grp = threadIdx.x / G;
    const int n = a.n, S = n | 1;
    extern __shared__ __align__(16) char smem[];
    if constexpr (!LIST) {
        const int slot = blockIdx.x * GPW + grp;
        const bool inslot = slot < a.slots;
        if (GPW == 1 && !inslot) return;   // the whole workgroup is past the last slot
        cx<T>* aglob = a.awork ? (cx<T>*)((char*)a.awork + a.awork_stride * (size_t)(inslot ? slot : 0)) : nullptr;
        SolveLds<T> L = solve_carve<T>(smem + grp * a.group_bytes, n, aglob);
        const int np = inslot ? a.npix[slot] : -1;
        const cx<Tg>* gam = (const cx<Tg>*)a.gamma + (size_t)(inslot ? slot : 0) * n * n;
        const long long pix = a.first + (inslot ? slot : 0);
        solve_generic_pixel<T, Tg, G, R, sizeof(T) == 4>(a, L, inslot, np, gam, pix);
    } else {
        static_assert(sizeof(T) == 8 && sizeof(Tg) == 8 && G == 128, "the re-solve runs in FP64 on two waves");
        constexpr int RG = R * 2;   // rows per lane of the 64-lane FP64 Gram
        char* area = (char*)a.awork + a.awork_stride * (size_t)blockIdx.x;
        cx<double>* gam = (cx<double>*)area;
        cx<double>* aglob = (cx<double>*)(area + (((size_t)n * n * sizeof(cx<double>) + 255) & ~(size_t)255));
        SolveLds<T> L = solve_carve<T>(smem, n, aglob);
        const size_t gbytes = gram_f64_group_bytes(n, a.g64.Nx, a.g64.Ny, kResolveCB);
        char* gscratch = smem + a.group_bytes + (threadIdx.x / 64) * gbytes;
        const int cnt = min(*a.list_count, a.list_cap);
        for (int e = blockIdx.x; e < cnt; e += gridDim.x) {
            const long long pix = a.list[e];
            const int np = gram_f64_pixel<64, RG, kResolveCB>(a.g64, true, pix, gam, gscratch);
            __threadfence();
            __syncthreads();
            solve_generic_pixel<T, Tg, G, R, false>(a, L, true, np, gam, pix);
            __syncthreads();
        }
    }
}

// Persistent launch above 64 dates: one pixel per two-wave workgroup per
// unit; each job's dynamic LDS fits the launch's (the largest of its jobs).
template <typename T, typename Tg, int R>
__global__ __launch_bounds__(128) void solve_generic_persistent(PersistArgs pa) {
    __shared__ int s_u;
    extern __shared__ __align__(16) char smem[];
    const int total = pa.unit_end[pa.njobs - 1];
    for (;;) {
        const int u0 = take_units(pa, total, &s_u);
        if (u0 >= total) break;
        const int u1 = min(u0 + pa.grab, total);
        for (int v = u0; v < u1; ++v) {
            const int u = total - 1 - v;
            const int j = job_of_unit(pa, u);
            const PJob& J = pa.job[j];
            const SolveArgs a = job_args(pa, J);
            const int slot = u - (j ? pa.unit_end[j - 1] : 0);
            cx<T>* aglob = a.awork ? (cx<T>*)((char*)a.awork + a.awork_stride * (size_t)slot) : nullptr;
            SolveLds<T> L = solve_carve<T>(smem, a.n, aglob);
            const cx<Tg>* gam = (const cx<Tg>*)a.gamma + (size_t)slot * a.n * a.n;
            solve_generic_pixel<T, Tg, 128, R, sizeof(T) == 4>(a, L, true, a.npix[slot], gam, a.first + slot);
            __syncthreads();
        }
    }
}

}  // namespace fringe
