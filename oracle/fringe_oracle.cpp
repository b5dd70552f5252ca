// fringe_oracle.cpp -- CPU restatement of fringe's phase-linking hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: the GPU library never
// links, loads or calls it.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use it.
//
// It restates, in plain C++ with an OpenMP pixel loop and FP64 LAPACK
// (zheevr / zpotrf / zpotri resolved at run time), the per-pixel semantics of
//   * src/nmap/nmap.cpp:204-233, :345-473  (amplitude, validity, sort, KS sweep)
//   * src/nmap/KS2sample.hpp:42-144        (KolmogorovProb, KS2sample::test)
//   * include/fringe/ulongmask.hpp:55-94   (bit layout)
//   * include/fringe/EigenLapack.hpp:131-251 (EVWorker)
//   * src/evd/evd.cpp:512-788              (EVD / MLE / STBAS chain)
//   * src/phase_link/phase_link.cpp:479-666 (EMI chain with EVD fallback)
// with the row-block scheduler (evd.cpp:399-443, :861-868) collapsed to its
// net effect: a sliding window truncated only at the image edges.
//
// nmap follows the race-free definition (each pixel evaluates its whole window
// and writes only its own bits and count), which is what the reference CPU
// sweep computes whenever its unsynchronised updates do not collide.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <omp.h>
#include <vector>

typedef std::complex<double> zd;
typedef std::complex<float> zf;

// --------------------------------------------------------------------------
// LAPACK, resolved at run time (Fortran ABI, as EigenLapack.hpp:28-58 binds it)
// --------------------------------------------------------------------------
typedef void (*zheevr_fn)(char*, char*, char*, int*, zd*, int*, double*, double*,
                          int*, int*, double*, int*, double*, zd*, int*, int*,
                          zd*, int*, double*, int*, int*, int*, int*);
typedef void (*zpotrf_fn)(char*, int*, zd*, int*, int*);
typedef void (*zpotri_fn)(char*, int*, zd*, int*, int*);

static zheevr_fn p_zheevr = nullptr;
static zpotrf_fn p_zpotrf = nullptr;
static zpotri_fn p_zpotri = nullptr;

extern "C" int oracle_lapack_load(const char* path_hint) {
    if (p_zheevr && p_zpotrf && p_zpotri) return 0;
    // MKL must stay single threaded inside the OpenMP pixel loop, as the
    // reference drivers force with OPENBLAS_NUM_THREADS=1 (evd.py:43).
    setenv("MKL_THREADING_LAYER", "SEQUENTIAL", 0);
    setenv("MKL_NUM_THREADS", "1", 0);
    const char* cands[] = {path_hint, getenv("FRINGE_ORACLE_LAPACK"),
                           "/opt/conda/lib/libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1",
                           "libmkl_rt.so", "liblapack.so.3", "libopenblas.so.0"};
    for (const char* c : cands) {
        if (!c || !*c) continue;
        void* h = dlopen(c, RTLD_NOW | RTLD_LOCAL);
        if (!h) continue;
        zheevr_fn a = (zheevr_fn)dlsym(h, "zheevr_");
        zpotrf_fn b = (zpotrf_fn)dlsym(h, "zpotrf_");
        zpotri_fn d = (zpotri_fn)dlsym(h, "zpotri_");
        if (a && b && d) {
            p_zheevr = a; p_zpotrf = b; p_zpotri = d;
            return 0;
        }
        dlclose(h);
    }
    return -1;
}

// Per-thread workspace with the reference's LAPACK settings
// (EigenLapack.hpp:111-128: LWORK=2N, LRWORK=24N, LIWORK=10N; :131-202: RANGE='I',
// UPLO='U', ABSTOL=1e-6).
struct EigWork {
    int n = 0;
    std::vector<double> w, rwork;
    std::vector<zd> z, work;
    std::vector<int> isuppz, iwork;
    void prepare(int nn) {
        n = nn;
        w.assign(n, 0.0);
        z.assign((size_t)n * n, zd(0, 0));
        isuppz.assign(2 * n, 0);
        work.assign(2 * n, zd(0, 0));
        rwork.assign(24 * n, 0.0);
        iwork.assign(10 * n, 0);
    }
    // index is 1-based (1 = smallest, n = largest); destroys A (column-major)
    int eigen_with_index(zd* A, int index, bool vec) {
        char jobz = vec ? 'V' : 'N', range = 'I', uplo = 'U';
        int N = n, lda = n, il = index, iu = index, m = 0, ldz = n, info = 0;
        double vl = 0, vu = 0, abstol = 1.0e-6;
        int lwork = 2 * n, lrwork = 24 * n, liwork = 10 * n;
        p_zheevr(&jobz, &range, &uplo, &N, A, &lda, &vl, &vu, &il, &iu, &abstol, &m,
                 w.data(), z.data(), &ldz, isuppz.data(), work.data(), &lwork,
                 rwork.data(), &lrwork, iwork.data(), &liwork, &info);
        return info;
    }
    // zpotrf + zpotri on the upper triangle, then lower := conj(upper)
    int pd_inverse(zd* A) {
        char uplo = 'U';
        int N = n, lda = n, info = 0;
        p_zpotrf(&uplo, &N, A, &lda, &info);
        if (info != 0) return info;
        p_zpotri(&uplo, &N, A, &lda, &info);
        for (int c = 0; c < n; ++c)
            for (int r = c + 1; r < n; ++r) A[(size_t)c * n + r] = std::conj(A[(size_t)r * n + c]);
        return info;
    }
};

// --------------------------------------------------------------------------
// SHP bit layout (ulongmask.hpp:55-94): flat = (ii+Ny)(2Nx+1) + (jj+Nx),
// word = flat/32, bit = flat%32, LSB first.
// --------------------------------------------------------------------------
static inline bool get_bit(const uint32_t* w, int ii, int jj, int Nx, int Ny) {
    int flat = (ii + Ny) * (2 * Nx + 1) + jj + Nx;
    return (w[flat >> 5] >> (flat & 31)) & 1u;
}
static inline void set_bit(uint32_t* w, int ii, int jj, int Nx, int Ny) {
    int flat = (ii + Ny) * (2 * Nx + 1) + jj + Nx;
    w[flat >> 5] |= (1u << (flat & 31));
}
extern "C" int oracle_mask_words(int Nx, int Ny) {
    return (int)std::ceil(((2 * Ny + 1) * (2 * Nx + 1)) / 32.0);
}

// --------------------------------------------------------------------------
// Kolmogorov distribution tail, ROOT's piecewise approximation as used by
// KS2sample::KolmogorovProb (KS2sample.hpp:42-77).  The constants are the
// ones that approximation is defined with: sqrt(2*pi) to 9 digits, -pi^2/8
// times 1, 9, 25, and the series exponents -2 j^2.
// --------------------------------------------------------------------------
extern "C" double oracle_kolmogorov_prob(double z) {
    const double w = 2.50662827;
    const double c1 = -1.2337005501361697, c2 = -11.103304951225528, c3 = -30.842513753404244;
    double u = std::fabs(z);
    if (u < 0.2) return 1.0;
    if (u < 0.755) {
        double v = 1.0 / (u * u);
        return 1.0 - w * (std::exp(c1 * v) + std::exp(c2 * v) + std::exp(c3 * v)) / u;
    }
    if (u < 6.8116) {
        double r[4] = {0, 0, 0, 0};
        double v = u * u;
        int maxj = std::max(1, (int)std::round(3.0 / u));
        for (int j = 0; j < maxj && j < 4; ++j) r[j] = std::exp(-2.0 * (j + 1) * (j + 1) * v);
        return 2.0 * (r[0] - r[1] + r[2] - r[3]);
    }
    return 0.0;
}

// Two-sample KS statistic on two ascending arrays of equal length n, walked
// the way KS2sample::test does (KS2sample.hpp:91-144): one element per step
// on strict order, every copy of a tied value from both sides on a tie, the
// running difference accumulated in +/- 1/n steps in double and sampled
// after each step while both sides still have elements.
extern "C" double oracle_ks2_statistic(const float* a, const float* b, int n) {
    const double s = 1.0 / (double)n;
    int ia = 0, ib = 0;
    double rdiff = 0.0, rdmax = 0.0;
    for (int i = 0; i < 2 * n; ++i) {
        if (a[ia] < b[ib]) {
            rdiff -= s;
            ++ia;
        } else if (a[ia] > b[ib]) {
            rdiff += s;
            ++ib;
        } else {
            float x = a[ia];
            while (ia < n && a[ia] == x) { rdiff -= s; ++ia; }
            while (ib < n && b[ib] == x) { rdiff += s; ++ib; }
        }
        if (ia >= n || ib >= n) break;
        rdmax = std::max(rdmax, std::fabs(rdiff));
    }
    return std::max(rdmax, std::fabs(rdiff));
}

extern "C" double oracle_ks2_pvalue(const float* a, const float* b, int n) {
    double rn = n;
    double d = oracle_ks2_statistic(a, b, n);
    return oracle_kolmogorov_prob(d * std::sqrt(rn * rn / (rn + rn)));
}

// --------------------------------------------------------------------------
// Two-sample Anderson-Darling test as AD2unique evaluates it
// (AD2unique.hpp:125-351): Scholz & Stephens' k-sample statistic A2_kN for
// k = 2 equal samples, standardised by its variance sigma_N
// (Scholz & Stephens 1987, eq. 4 with the harmonic sums h and g), and a
// p-value by linear interpolation of the log-odds in the k = 2 column of
// their quantile table.
// --------------------------------------------------------------------------
// quantiles t(p) of the standardised statistic for k = 2 (AD2unique.hpp:8-52,
// every 8th entry) and the upper-tail probabilities they belong to
static const double kAdT[35] = {
    -1.1954, -1.1786, -1.166, -1.1407, -1.1253, -1.0777, -1.0489, -0.9978, -0.9417, -0.8981,
    -0.8598, -0.7258, -0.5966, -0.4572, -0.2966, -0.1009, 0.1571, 0.5357, 1.2255, 1.5262,
    1.9633, 2.7314, 3.7825, 4.1241, 4.6044, 5.409, 6.4954, 6.8279, 7.2755, 8.1885,
    9.3061, 9.6132, 10.0989, 10.8825, 11.8537};
static const double kAdP[35] = {
    .00001, .00005, .0001, .0005, .001, .005, .01, .025, .05, .075, .1, .2, .3, .4, .5, .6, .7, .8, .9,
    .925, .95, .975, .99, .9925, .995, .9975, .999, .99925, .9995, .99975, .9999, .999925,
    .99995, .999975, .99999};

extern "C" double oracle_ad2_sigma(int n1, int n2) {
    const int N = n1 + n2;
    const double H = 1.0 / (1.0 * n1) + 1.0 / (1.0 * n2);
    double h = 0.0, g = 0.0;
    std::vector<double> inv(N > 0 ? N : 1, 0.0);
    if (N < 2000) {
        for (int i = 1; i < N; ++i) {
            inv[i] = 1.0 / i;
            h += inv[i];
        }
        for (int i = 1; i < N - 1; ++i)
            for (int j = i + 1; j < N; ++j) g += inv[N - i] * inv[j];
    } else {
        h = std::log(double(N - 1)) + 0.5772156649015328606065120900824024;
        g = M_PI * M_PI / 6.0;
    }
    const double k = 2.0, k2 = std::pow(k, 2);
    const double a = (4 * g - 6) * (k - 1) + (10 - 6 * g) * H;
    const double b = (2 * g - 4) * k2 + 8 * h * k + (2 * g - 14 * h - 4) * H - 8 * h + 4 * g - 6;
    const double c = (6 * h + 2 * g - 2) * k2 + (4 * h - 4 * g + 6) * k + (2 * h - 6) * H + 4 * h;
    const double d = (2 * h + 6) * k2 - 4 * h * k;
    double s = 0.0;
    s += a * std::pow(double(N), 3) + b * std::pow(double(N), 2) + c * N + d;
    s /= (double(N - 1) * double(N - 2) * double(N - 3));
    return std::sqrt(s);
}

extern "C" double oracle_ad2_pvalue_of(double t) {
    int lo = 0;
    int first = 35;   // first table entry >= t
    for (int i = 0; i < 35; ++i)
        if (t <= kAdT[i]) {
            first = i;
            break;
        }
    lo = first - 1;
    if (first == 35) lo = 33;   // beyond the table: last segment
    if (lo < 0) lo = 0;         // before the table: first segment
    const int hi = lo + 1;
    const double l1 = std::log((1.0 - kAdP[lo]) / kAdP[lo]), l2 = std::log((1.0 - kAdP[hi]) / kAdP[hi]);
    const double l0 = (l1 - l2) * (t - kAdT[hi]) / (kAdT[lo] - kAdT[hi]) + l2;
    return std::exp(l0) / (1. + std::exp(l0));
}

// a, b ascending, n each; ties are merged b first.  Only the merge pattern
// enters: with m_j = number of a among the first j+1 merged values, the
// inner sums of both samples equal S = sum_{j < 2n-1} (2n m_j - n (j+1))^2 /
// ((j+1)(2n - j - 1)), accumulated in ascending j.
extern "C" double oracle_ad2_statistic(const float* a, const float* b, int n) {
    const int l = 2 * n;
    const double nsum = l;
    int ia = 0, ib = 0, m = 0;
    double S = 0.0;
    for (int j = 0; j < l - 1; ++j) {
        const bool take_a = (ia < n && ib < n) ? (a[ia] < b[ib]) : (ia < n);
        if (take_a) {
            ++ia;
            ++m;
        } else {
            ++ib;
        }
        const double bj = j + 1;
        const double t = nsum * m - (double)n * bj;
        S = S + t * t / (bj * (nsum - bj));
    }
    double A = 0.0;
    A = A + S / (n * 1.0);
    A = A + S / (n * 1.0);
    A = A / nsum;
    return (A - 1) / oracle_ad2_sigma(n, n);
}

extern "C" double oracle_ad2_pvalue(const float* a, const float* b, int n) {
    return oracle_ad2_pvalue_of(oracle_ad2_statistic(a, b, n));
}

// --------------------------------------------------------------------------
// nmap: race-free full-window sweep (nmap.cpp:204-233, :345-473;
// nmap_cuda.cu:259-334 semantics).  method 0 = KS2, 1 = AD2.  The reference
// tests a pair (p, q) with p first when q lies in the half-plane it sweeps
// (row offset > 0, or = 0 with column offset > 0), so the other half of a
// pixel's window is tested with the neighbour first (AD2 merges ties
// asymmetrically; KS2 is symmetric).
//   slc   : [nbands][rows][cols] complex64
//   mask  : [rows][cols] uint8 or NULL
//   aconst: [nbands] amplitudeConstant per band, <= 0 meaning "missing"
//   wts   : [rows][cols][nulong] (zeroed here), count: [rows][cols]
// --------------------------------------------------------------------------
extern "C" int oracle_nmap(const float* slc, const uint8_t* mask, int rows, int cols,
                           int nbands, const double* aconst, int Nx, int Ny, double prob,
                           int method, uint32_t* wts, int32_t* count, int nthreads) {
    const int nulong = oracle_mask_words(Nx, Ny);
    const size_t npx = (size_t)rows * cols;
    std::vector<double> alpha(nbands, 1.0);
    for (int b = 0; b < nbands; ++b) {
        double c = aconst ? aconst[b] : 0.0;
        alpha[b] = (c <= 0.0) ? 1.0 : c;
    }
    const double norm = alpha[0];
    for (int b = 0; b < nbands; ++b) alpha[b] /= norm;
    alpha[0] = 1.0;

    std::vector<float> amp(npx * nbands, 0.0f);
    std::vector<uint8_t> valid(npx, 1);
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (long long p = 0; p < (long long)npx; ++p) {
        uint8_t ok = mask ? (mask[p] != 0) : 1;
        for (int b = 0; b < nbands; ++b) {
            const float* z = slc + 2 * ((size_t)b * npx + p);
            float val = std::abs(zf(z[0], z[1])) / alpha[b];
            ok = ok && (val != 0.) && !std::isnan(val);
            if (ok) amp[p * nbands + b] = val;
        }
        valid[p] = ok;
        if (ok) std::sort(&amp[p * nbands], &amp[p * nbands] + nbands);
    }
    std::memset(wts, 0, sizeof(uint32_t) * npx * nulong);
    std::memset(count, 0, sizeof(int32_t) * npx);
#pragma omp parallel for schedule(dynamic, 16)
    for (int r = 0; r < rows; ++r) {
        for (int c = 0; c < cols; ++c) {
            size_t p = (size_t)r * cols + c;
            if (!valid[p]) continue;
            uint32_t* w = wts + p * nulong;
            int cnt = 0;
            for (int ii = -Ny; ii <= Ny; ++ii) {
                int rr = r + ii;
                if (rr < 0 || rr >= rows) continue;
                for (int jj = -Nx; jj <= Nx; ++jj) {
                    int cc = c + jj;
                    if (cc < 0 || cc >= cols) continue;
                    size_t q = (size_t)rr * cols + cc;
                    if (!valid[q]) continue;
                    if (q == p) {
                        set_bit(w, 0, 0, Nx, Ny);
                        ++cnt;
                    } else {
                        const bool fwd = ii > 0 || (ii == 0 && jj > 0);
                        const float* x = &amp[(fwd ? p : q) * nbands];
                        const float* y = &amp[(fwd ? q : p) * nbands];
                        const double pv = method == 1 ? oracle_ad2_pvalue(x, y, nbands)
                                                      : oracle_ks2_pvalue(x, y, nbands);
                        if (!(pv >= prob)) continue;
                        set_bit(w, ii, jj, Nx, Ny);
                        ++cnt;
                    }
                }
            }
            count[p] = cnt;
        }
    }
    return 0;
}

// --------------------------------------------------------------------------
// calamp (calamp.cpp:180-262): per band, the mean of |z| (rounded to float as
// std::abs of complex<float>, summed in double) over pixels with |z| != 0,
// |z| not NaN and mask > 0; 1.0 for a band without valid pixels.
// --------------------------------------------------------------------------
extern "C" int oracle_calamp(const float* slc, const uint8_t* mask, int rows, int cols, int nbands,
                             double* out) {
    const size_t npx = (size_t)rows * cols;
    for (int b = 0; b < nbands; ++b) {
        double sum = 0.0, norm = 0.0;
        for (size_t p = 0; p < npx; ++p) {
            const float* z = slc + 2 * ((size_t)b * npx + p);
            double a = std::abs(zf(z[0], z[1]));
            a = std::isnan(a) ? 0.0 : a;
            const int valid = (a != 0.0) * ((mask ? mask[p] : 1) > 0);
            sum += valid * a;
            norm += valid;
        }
        out[b] = norm == 0 ? 1.0 : sum / norm;
    }
    return 0;
}

// --------------------------------------------------------------------------
// despeck: SHP-masked neighbourhood mean (despeck.cpp:300-433) over a whole
// image (the reference's blocks only bound memory: windows are truncated at
// the image edges, as here).  mode 0: |b1|; 1: b1 conj(b2); 2: b1 conj(b2)
// normalised to coherence.  out: [rows][cols] complex64 (zeros where the
// centre bit is clear); the amplitude output is its real part.
// --------------------------------------------------------------------------
extern "C" int oracle_despeck(const float* b1, const float* b2, const uint32_t* wts, int rows, int cols,
                              int Nx, int Ny, int mode, float* out) {
    const int nulong = oracle_mask_words(Nx, Ny);
    const size_t npx = (size_t)rows * cols;
    std::vector<zf> c1(npx), c2(npx);
    for (size_t p = 0; p < npx; ++p) {
        const zf a(b1[2 * p], b1[2 * p + 1]);
        if (mode == 0) {
            c1[p] = zf(std::abs(a), 0.0f);
            c2[p] = zf(1.0f, 0.0f);
            continue;
        }
        const zf b(b2[2 * p], b2[2 * p + 1]);
        const float m1 = std::abs(a), m2 = std::abs(b);
        c1[p] = a * std::conj(b);
        c2[p] = (mode == 2) ? zf(m1 * m1, m2 * m2) : zf(1.0f, 0.0f);
    }
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) {
            const size_t p = (size_t)r * cols + c;
            const uint32_t* w = wts + p * nulong;
            zf res(0.0f, 0.0f);
            auto bit = [&](int ii, int jj) {
                const int f = (ii + Ny) * (2 * Nx + 1) + jj + Nx;
                return (w[f >> 5] >> (f & 31)) & 1u;
            };
            if (bit(0, 0)) {
                zf val(0.0f, 0.0f), sumw(0.0f, 0.0f);
                for (int ii = std::max(r - Ny, 0); ii <= std::min(rows - 1, r + Ny); ++ii)
                    for (int jj = std::max(c - Nx, 0); jj <= std::min(cols - 1, c + Nx); ++jj)
                        if (bit(ii - r, jj - c)) {
                            val += c1[(size_t)ii * cols + jj];
                            sumw += c2[(size_t)ii * cols + jj];
                        }
                if (sumw.real() > 0) {
                    if (mode == 2) {
                        if (sumw.imag() > 0) res = val / (std::sqrt(sumw.real()) * std::sqrt(sumw.imag()));
                    } else {
                        res = val / sumw.real();
                    }
                }
            }
            out[2 * p] = res.real();
            out[2 * p + 1] = res.imag();
        }
    return 0;
}

// --------------------------------------------------------------------------
// Phase linking.  module: 0 = evd (evd.cpp), 1 = phase_link (phase_link.cpp).
// method: 0 = MLE, 1 = EVD, 2 = STBAS (phase_link always runs EMI; STBAS
// there only limits the temporal-coherence sum, phase_link.cpp:654).
//   slc  : [nbands][rows][cols] complex64, wts: [rows][cols][nulong]
//   phase: [nbands][rows][cols] complex64, tcorr: [rows][cols], comp: [rows][cols] complex64
// Outputs are written only for processed pixels in rows [row0, row1); the
// caller zero-fills them.  gap (optional, [rows][cols]) receives the distance
// from the selected eigenvalue to its nearest neighbour divided by the
// spread of the spectrum (diagnostic for near-degenerate pixels), or -1.
// vmin (optional, [rows][cols]) receives sqrt(n) min_t |v_t| of the selected
// unit eigenvector: the smallest component relative to the RMS one, whose
// phase an eigenvector error of absolute size d fixes only to ~d / |v_t|.
// Returns 0, or 101 for an invalid STBAS bandwidth (evd.cpp:74-91).
// --------------------------------------------------------------------------
enum { MOD_EVD = 0, MOD_PHASE_LINK = 1 };
enum { METH_MLE = 0, METH_EVD = 1, METH_STBAS = 2 };

extern "C" int oracle_link(const float* slc, const uint32_t* wts, int rows, int cols,
                           int nbands, int Nx, int Ny, int module, int method,
                           int minNeighbors, int miniStackCount, int bandWidth, int row0,
                           int row1, float* phase, float* tcorr, float* comp, double* gap,
                           double* vmin, int nthreads) {
    if (oracle_lapack_load(nullptr) != 0) return -100;
    const int n = nbands;
    const bool isstbas = (method == METH_STBAS);
    const bool ismle = (method == METH_MLE);
    const int BW = bandWidth;
    if (isstbas && (BW <= 0 || BW >= n - 1)) return 101;
    const int nulong = oracle_mask_words(Nx, Ny);
    const size_t npx = (size_t)rows * cols;
    const int ref = miniStackCount - 1;
    const int minN = (module == MOD_EVD) ? 2 : minNeighbors;
    if (nthreads > 0) omp_set_num_threads(nthreads);

#pragma omp parallel
    {
        EigWork ew;
        ew.prepare(n);
        std::vector<zd> numer((size_t)n * n), covar((size_t)n * n), amp1((size_t)n * n),
            amp2((size_t)n * n), work((size_t)n * n);
        std::vector<double> asum(n);
        std::vector<zf> row(n), out(n);

#pragma omp for schedule(dynamic, 64)
        for (long long pp = (long long)row0 * cols; pp < (long long)row1 * cols; ++pp) {
            const uint32_t* cen = wts + (size_t)pp * nulong;
            if (!get_bit(cen, 0, 0, Nx, Ny)) continue;
            const int ci = (int)(pp / cols), cj = (int)(pp % cols);
            const int xmin = std::max(cj - Nx, 0), xmax = std::min(cols - 1, cj + Nx);
            const int ymin = std::max(ci - Ny, 0), ymax = std::min(rows - 1, ci + Ny);
            std::fill(numer.begin(), numer.end(), zd(0, 0));
            std::fill(asum.begin(), asum.end(), 0.0);
            int npix = 0;
            // Masked-neighbourhood Gram (evd.cpp:539-564): each product is a
            // complex<float> product, accumulated in complex<double>; the
            // amplitudes are float |z| squared in double.
            for (int ii = ymin; ii <= ymax; ++ii)
                for (int jj = xmin; jj <= xmax; ++jj) {
                    if (!get_bit(cen, ii - ci, jj - cj, Nx, Ny)) continue;
                    ++npix;
                    size_t q = (size_t)ii * cols + jj;
                    for (int t = 0; t < n; ++t) {
                        const float* z = slc + 2 * ((size_t)t * npx + q);
                        row[t] = zf(z[0], z[1]);
                    }
                    for (int ti = 0; ti < n; ++ti) {
                        double a = std::pow((double)std::abs(row[ti]), 2);
                        asum[ti] += a;
                        for (int tj = ti + 1; tj < n; ++tj) {
                            zf prod = row[ti] * std::conj(row[tj]);
                            numer[(size_t)tj * n + ti] += zd(prod.real(), prod.imag());
                        }
                    }
                }
            if (npix < minN) continue;

            // Coherence (evd.cpp:569-582): column-major, element (ti,tj) at tj*n+ti.
            for (int ti = 0; ti < n; ++ti) {
                for (int tj = ti + 1; tj < n; ++tj) {
                    zd res = numer[(size_t)tj * n + ti] / std::sqrt(asum[ti] * asum[tj]);
                    covar[(size_t)tj * n + ti] = res;
                    covar[(size_t)ti * n + tj] = std::conj(res);
                }
                covar[(size_t)ti * n + ti] = zd(1.0, 0.0);
            }

            int sel_index = -1;  // eigen index (1-based) of the solution, for gap
            std::vector<zd>* sel_matrix = nullptr;
            bool ok = true;
            float code = 0.0f;

            if (module == MOD_EVD && ismle) {
                // evd.cpp:584-688
                for (int ti = 0; ti < n; ++ti) {
                    for (int tj = ti + 1; tj < n; ++tj) {
                        double coh = std::abs(covar[(size_t)tj * n + ti]);
                        amp1[(size_t)tj * n + ti] = zd(coh, 0);
                        amp1[(size_t)ti * n + tj] = zd(coh, 0);
                    }
                    amp1[(size_t)ti * n + ti] = zd(1, 0);
                }
                amp2 = covar;
                int st = ew.eigen_with_index(amp2.data(), 1, false);
                if (st != 0) { code = -1; ok = false; }
                else if (ew.w[0] < 1.0e-6) { code = -2; ok = false; }
                if (ok) {
                    amp2 = amp1;
                    st = ew.eigen_with_index(amp2.data(), 1, false);
                    if (st != 0) { code = -3; ok = false; }
                    else if (ew.w[0] < 1.0e-6) { code = -4; ok = false; }
                }
                if (ok) {
                    amp2 = amp1;
                    if (ew.pd_inverse(amp2.data()) != 0) { code = -5; ok = false; }
                }
                if (ok) {
                    for (size_t k = 0; k < amp2.size(); ++k) amp2[k] *= covar[k];
                    work = amp2;
                    st = ew.eigen_with_index(amp2.data(), 1, true);
                    if (st != 0) { code = -6; ok = false; }
                    else if (ew.w[0] < 1.0e-6) { code = -7; ok = false; }
                    sel_index = 1;
                    sel_matrix = &work;
                }
            } else if (module == MOD_EVD) {
                // EVD / STBAS (evd.cpp:689-732)
                if (isstbas) {
                    for (int ti = 0; ti < n; ++ti)
                        for (int tj = ti + BW + 1; tj < n; ++tj) {
                            covar[(size_t)ti * n + tj] = zd(0, 0);
                            covar[(size_t)tj * n + ti] = zd(0, 0);
                        }
                }
                amp1 = covar;
                work = covar;
                int st = ew.eigen_with_index(amp1.data(), n, true);
                if (st != 0) { code = -6; ok = false; }
                else if (ew.w[0] < 1.0e-6) { code = -7; ok = false; }
                sel_index = n;
                sel_matrix = &work;
            } else {
                // phase_link EMI with EVD fallback (phase_link.cpp:540-618)
                amp1 = covar;
                int st = ew.eigen_with_index(amp1.data(), 1, false);
                if (st != 0) { code = -1; ok = false; }
                bool run_evd = false;
                if (ok) {
                    for (int ti = 0; ti < n; ++ti) {
                        for (int tj = ti + 1; tj < n; ++tj) {
                            double coh = std::abs(covar[(size_t)tj * n + ti]);
                            amp1[(size_t)tj * n + ti] = zd(coh, 0);
                            amp1[(size_t)ti * n + tj] = zd(coh, 0);
                        }
                        amp1[(size_t)ti * n + ti] = zd(1, 0);
                    }
                    amp2 = amp1;
                    if (ew.pd_inverse(amp2.data()) != 0) run_evd = true;
                    if (!run_evd) {
                        for (size_t k = 0; k < amp2.size(); ++k) amp2[k] *= covar[k];
                        work = amp2;
                        if (ew.eigen_with_index(amp2.data(), 1, true) != 0) run_evd = true;
                        sel_index = 1;
                        sel_matrix = &work;
                    }
                    if (run_evd) {
                        amp1 = covar;
                        work = covar;
                        if (ew.eigen_with_index(amp1.data(), n, true) != 0) { code = -6; ok = false; }
                        sel_index = n;
                        sel_matrix = &work;
                    }
                }
            }
            if (!ok) {
                tcorr[pp] = code;
                continue;
            }

            // Reference phase and output quantisation (evd.cpp:738-749)
            {
                double ph0 = std::arg(ew.z[ref]);
                for (int t = 0; t < n; ++t) {
                    double res = std::arg(ew.z[t]);
                    zd e = std::exp(zd(0.0, 1.0) * (res - ph0));
                    out[t] = zf((float)e.real(), (float)e.imag());
                }
                out[ref] = zf(1.0f, 0.0f);
            }
            // Compressed SLC from the centre pixel (evd.cpp:755-762)
            {
                zd sum(0, 0);
                for (int t = ref; t < n; ++t) {
                    const float* z = slc + 2 * ((size_t)t * npx + pp);
                    zf prod = zf(z[0], z[1]) * std::conj(out[t]);
                    sum += zd(prod.real(), prod.imag());
                }
                zd cval = sum / (1.0 * (n - ref));
                comp[2 * pp] = (float)cval.real();
                comp[2 * pp + 1] = (float)cval.imag();
            }
            // Temporal coherence (evd.cpp:770-786)
            {
                zd tc(0, 0);
                int counter = 0;
                for (int ti = 0; ti < n; ++ti) {
                    int ulim = isstbas ? std::min(n, ti + BW + 1) : n;
                    for (int tj = ti + 1; tj < ulim; ++tj) {
                        double a = std::arg(covar[(size_t)tj * n + ti]) - std::arg(out[ti]) +
                                   std::arg(out[tj]);
                        tc += std::exp(zd(0.0, 1.0) * a);
                        ++counter;
                    }
                }
                tcorr[pp] = (float)(std::abs(tc) / (counter * 1.0));
            }
            for (int t = 0; t < n; ++t) {
                phase[2 * ((size_t)t * npx + pp)] = out[t].real();
                phase[2 * ((size_t)t * npx + pp) + 1] = out[t].imag();
            }

            if (vmin) {
                double mn = 1e300;
                for (int t = 0; t < n; ++t) mn = std::min(mn, std::abs(ew.z[t]));
                vmin[pp] = mn * std::sqrt((double)n);
            }
            if (gap && sel_matrix) {
                // Diagnostic: neighbour gap of the selected eigenvalue over the
                // spectral spread, from the full FP64 spectrum.
                // (eigenvalues 1, 2, n-1 and n are all it needs)
                std::vector<zd> m = *sel_matrix;
                EigWork all;
                all.prepare(n);
                auto eig = [&](int k) {
                    std::vector<zd> mm = m;
                    all.eigen_with_index(mm.data(), k, false);
                    return all.w[0];
                };
                if (n < 2) {
                    gap[pp] = 1.0;
                    continue;
                }
                const double l1 = eig(1), ln = eig(n);
                const double spread = std::max(ln - l1, 1e-300);
                const double g = (sel_index == 1) ? eig(2) - l1 : ln - eig(n - 1);
                gap[pp] = g / spread;
            }
        }
    }
    return 0;
}

// Single Hermitian eigenpair through the same LAPACK call the reference makes
// (EigenLapack.hpp:131-215).  A is column-major n x n and is preserved.
extern "C" int oracle_eigen_index(const double* A_interleaved, int n, int index, double* eigval,
                                  double* eigvec_interleaved) {
    if (oracle_lapack_load(nullptr) != 0) return -100;
    EigWork ew;
    ew.prepare(n);
    std::vector<zd> a((size_t)n * n);
    for (size_t k = 0; k < a.size(); ++k) a[k] = zd(A_interleaved[2 * k], A_interleaved[2 * k + 1]);
    int info = ew.eigen_with_index(a.data(), index, eigvec_interleaved != nullptr);
    *eigval = ew.w[0];
    if (eigvec_interleaved)
        for (int k = 0; k < n; ++k) {
            eigvec_interleaved[2 * k] = ew.z[k].real();
            eigvec_interleaved[2 * k + 1] = ew.z[k].imag();
        }
    return info;
}

// PD inverse through zpotrf + zpotri (EigenLapack.hpp:217-251); in place.
extern "C" int oracle_pd_inverse(double* A_interleaved, int n) {
    if (oracle_lapack_load(nullptr) != 0) return -100;
    EigWork ew;
    ew.prepare(n);
    std::vector<zd> a((size_t)n * n);
    for (size_t k = 0; k < a.size(); ++k) a[k] = zd(A_interleaved[2 * k], A_interleaved[2 * k + 1]);
    int info = ew.pd_inverse(a.data());
    for (size_t k = 0; k < a.size(); ++k) {
        A_interleaved[2 * k] = a[k].real();
        A_interleaved[2 * k + 1] = a[k].imag();
    }
    return info;
}
