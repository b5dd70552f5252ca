// engine.hip -- host side of the tile engine: argument checks, workspace
// layout, the KS decision table and the C ABI entry points declared in
// include/fringe_amd.h.
#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "engine_internal.hpp"
#include "link_solve.hpp"

namespace fringe {

static thread_local std::string g_last_error;

void set_last_error(const char* what, hipError_t e) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
}
void set_last_error(const std::string& what) { g_last_error = what; }

int device_lds_limit() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    static int cached[64] = {0};
    if (dev >= 0 && dev < 64 && cached[dev]) return cached[dev];
    hipDeviceProp_t prop;
    int lim = 65536;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.sharedMemPerBlock > 0)
        lim = (int)prop.sharedMemPerBlock;
    if (dev >= 0 && dev < 64) cached[dev] = lim;
    return lim;
}

// Kolmogorov distribution tail as KS2sample::KolmogorovProb evaluates it
// (KS2sample.hpp:42-77): ROOT's piecewise approximation with sqrt(2 pi)
// to 9 digits, -pi^2/8 * {1, 9, 25} and the series exponents -2 j^2.
static double kolmogorov_prob(double z) {
    const double w = 2.50662827;
    const double c1 = -1.2337005501361697, c2 = -11.103304951225528, c3 = -30.842513753404244;
    const double u = std::fabs(z);
    if (u < 0.2) return 1.0;
    if (u < 0.755) {
        const double v = 1.0 / (u * u);
        return 1.0 - w * (std::exp(c1 * v) + std::exp(c2 * v) + std::exp(c3 * v)) / u;
    }
    if (u < 6.8116) {
        double r[4] = {0.0, 0.0, 0.0, 0.0};
        const double v = u * u;
        const int maxj = std::max(1, (int)std::round(3.0 / u));
        for (int j = 0; j < maxj && j < 4; ++j) r[j] = std::exp(-2.0 * (j + 1) * (j + 1) * v);
        return 2.0 * (r[0] - r[1] + r[2] - r[3]);
    }
    return 0.0;
}


// ---------------------------------------------------------------------------
// Two-sample Anderson-Darling decision (AD2unique.hpp): sigma_N of Scholz &
// Stephens' A2_kN for k = 2 (AD2unique.hpp:162-203) and the p-value by
// log-odds interpolation in the k = 2 column of their quantile table
// (AD2unique.hpp:8-52, :125-160).
// ---------------------------------------------------------------------------
static const double kAdQuantile[35] = {
    -1.1954, -1.1786, -1.166, -1.1407, -1.1253, -1.0777, -1.0489, -0.9978, -0.9417, -0.8981,
    -0.8598, -0.7258, -0.5966, -0.4572, -0.2966, -0.1009, 0.1571, 0.5357, 1.2255, 1.5262,
    1.9633, 2.7314, 3.7825, 4.1241, 4.6044, 5.409, 6.4954, 6.8279, 7.2755, 8.1885,
    9.3061, 9.6132, 10.0989, 10.8825, 11.8537};
static const double kAdProb[35] = {
    .00001, .00005, .0001, .0005, .001, .005, .01, .025, .05, .075, .1, .2, .3, .4, .5, .6, .7, .8, .9,
    .925, .95, .975, .99, .9925, .995, .9975, .999, .99925, .9995, .99975, .9999, .999925,
    .99995, .999975, .99999};

static double ad2_sigma(int n1, int n2) {
    const int N = n1 + n2;
    const double H = 1.0 / (1.0 * n1) + 1.0 / (1.0 * n2);
    double h = 0.0, g = 0.0;
    if (N < 2000) {
        std::vector<double> inv(N, 0.0);
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
    double v = 0.0;
    v += a * std::pow(double(N), 3) + b * std::pow(double(N), 2) + c * N + d;
    v /= (double(N - 1) * double(N - 2) * double(N - 3));
    return std::sqrt(v);
}

static double ad2_pvalue(double t) {
    int first = 35;
    for (int i = 0; i < 35; ++i)
        if (t <= kAdQuantile[i]) {
            first = i;
            break;
        }
    int lo = (first == 35) ? 33 : first - 1;
    if (lo < 0) lo = 0;
    const int hi = lo + 1;
    const double l1 = std::log((1.0 - kAdProb[lo]) / kAdProb[lo]);
    const double l2 = std::log((1.0 - kAdProb[hi]) / kAdProb[hi]);
    const double l0 = (l1 - l2) * (t - kAdQuantile[hi]) / (kAdQuantile[lo] - kAdQuantile[hi]) + l2;
    return std::exp(l0) / (1. + std::exp(l0));
}

// Doubles mapped to unsigned keys in the same order (finite values).
static uint64_t dkey(double x) {
    uint64_t b;
    std::memcpy(&b, &x, 8);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
static double dval(uint64_t k) {
    const uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double x;
    std::memcpy(&x, &b, 8);
    return x;
}

// Largest standardised statistic t with p(t) >= prob (p decreases with t);
// -inf when no statistic is similar, +inf when every one is.  The search
// starts at -8, where p is still 1: the interpolated log-odds overflow exp
// (and p becomes NaN) only below about -8.5, and no statistic goes below
// -1/sigma_N >= -2.2 (A2_kN >= 0).
static double ad2_threshold(double prob) {
    const double lo0 = -8.0, hi0 = 1.0e6;
    if (!(ad2_pvalue(lo0) >= prob)) return -INFINITY;
    if (ad2_pvalue(hi0) >= prob) return INFINITY;
    uint64_t lo = dkey(lo0), hi = dkey(hi0);   // p(lo) >= prob > p(hi)
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (ad2_pvalue(dval(mid)) >= prob) lo = mid;
        else hi = mid;
    }
    return dval(lo);
}


// ---------------------------------------------------------------------------
// phase linking
// ---------------------------------------------------------------------------
int link_check(const fringe_link_tile* t) {
    if (!t) return FRINGE_ERR_ARGUMENT;
    if (t->rows <= 0 || t->cols <= 0 || t->nbands <= 0 || t->Nx < 0 || t->Ny < 0)
        return FRINGE_ERR_ARGUMENT;
    if (t->nbands > kMaxBandsLink) return FRINGE_ERR_UNSUPPORTED;
    if (shp_words(t->Nx, t->Ny) > kMaxMaskWords) return FRINGE_ERR_UNSUPPORTED;
    if (t->row_begin < 0 || t->row_end > t->rows || t->row_begin > t->row_end) return FRINGE_ERR_ARGUMENT;
    if (t->ref_index < 0 || t->ref_index >= t->nbands) return FRINGE_ERR_ARGUMENT;
    if (t->module != FRINGE_MODULE_EVD && t->module != FRINGE_MODULE_PHASE_LINK) return FRINGE_ERR_ARGUMENT;
    if (t->method < FRINGE_METHOD_MLE || t->method > FRINGE_METHOD_STBAS) return FRINGE_ERR_ARGUMENT;
    if (t->precision < FRINGE_PREC_FP64 || t->precision > FRINGE_PREC_MIXED_BF16) return FRINGE_ERR_ARGUMENT;
    if (t->method == FRINGE_METHOD_STBAS && (t->bandwidth <= 0 || t->bandwidth >= t->nbands - 1))
        return FRINGE_ERR_BANDWIDTH;   // evd.cpp:74-91
    if ((long long)t->rows * t->cols >= (1ll << 31) - 1) return FRINGE_ERR_UNSUPPORTED;
    // packed records are addressed with 32-bit word offsets up to 64 dates
    if (link_uses_mfma(t->nbands, t->precision) && t->nbands <= 64 &&
        ((long long)t->rows * t->cols + 1) * 2 * t->nbands >= (1ll << 32))
        return FRINGE_ERR_UNSUPPORTED;
    // LDS footprint of the generic solver used above 64 dates (one pixel per
    // two-wave workgroup): the pixel's matrices plus the kernel's static LDS
    // LDS of the solver kernels: above 64 dates the work matrix moves to
    // global memory when it does not fit, the rest of a pixel's group must
    const size_t lds = (size_t)device_lds_limit();
    if (t->nbands <= 64) {
        if (link_certifies(t->nbands, t->precision) &&
            (rows_solver_selected(t->nbands, FRINGE_PREC_FP64) ? resolve_rows_lds_bytes(t->nbands, t->Nx, t->Ny)
                                                               : resolve_lds_bytes(t->nbands, t->Nx, t->Ny)) > lds)
            return FRINGE_ERR_UNSUPPORTED;
    } else {
        const size_t rest = (t->precision == FRINGE_PREC_FP64 ? solve_group_bytes<double>(t->nbands, false)
                                                              : solve_group_bytes<float>(t->nbands, false)) +
                            solve_static_lds(t->precision);
        if (rest > lds) return FRINGE_ERR_UNSUPPORTED;
        if (link_certifies(t->nbands, t->precision)) {
            const size_t r = rows_solver_selected(t->nbands, FRINGE_PREC_FP64)
                                 ? resolve_rows_lds_bytes(t->nbands, t->Nx, t->Ny)
                                 : resolve_generic_lds_bytes(t->nbands, t->Nx, t->Ny);
            if (r > lds) return FRINGE_ERR_UNSUPPORTED;
        }
    }
    return 0;
}

static int link_chunk_pixels(const fringe_link_tile* t) {
    const size_t per = link_gamma_bytes_per_pixel(t->nbands, t->precision);
    // Gamma of one round: ~2 GB, so that one solver launch holds tens of
    // generations of resident waves (their Gamma reads and launch tails then
    // overlap other waves' work); smaller rounds that keep Gamma in the
    // 256 MB last-level cache were measured slower
    const size_t target = (size_t)2048 << 20;
    long long rows = (long long)(target / (per * (size_t)t->cols));
    if (rows < 1) rows = 1;
    // whole pixel tiles of the staged Gram kernel
    const int th = link_uses_mfma(t->nbands, t->precision) ? gram_stage_rows(t->nbands, t->Nx, t->Ny) : 1;
    if (rows > th) rows -= rows % th;
    const long long out_rows = t->row_end - t->row_begin;
    if (rows > out_rows) rows = out_rows > 0 ? out_rows : 1;
    return (int)(rows * t->cols);
}

struct LinkLayout {
    size_t gamma, npix, pack, cert, queue, awork, awork_stride, rwork, total;
    int resolve_blocks;
};

static LinkLayout link_layout(const fringe_link_tile* t, int chunk, size_t area_budget) {
    const int n = t->nbands;
    const size_t e = link_uses_mfma(n, t->precision) ? 8 : 16;
    LinkLayout L{};
    size_t b = 0;
    L.cert = b;   // first, so its place does not depend on the chunk
    if (link_certifies(n, t->precision))   // the count, then the list of the tile's output pixels
        b += 256 + align256((size_t)(t->row_end - t->row_begin) * t->cols * sizeof(int));
    L.queue = b;
    b += 256;
    L.gamma = b;
    b += align256((size_t)chunk * n * n * e);
    L.npix = b;
    b += align256((size_t)chunk * sizeof(int));
    L.pack = b;
    if (link_uses_mfma(n, t->precision)) b += align256(mfma_pack_bytes(t->rows, t->cols, n));
    L.awork = b;
    if (n > 64 && !solve_a_in_lds(n, t->precision)) {   // the solver's work matrices, one per slot
        const size_t el = t->precision == FRINGE_PREC_FP64 ? 16 : 8;
        L.awork_stride = align256((size_t)n * (n | 1) * el);
        b += (size_t)chunk * L.awork_stride;
    }
    L.rwork = b;   // the FP64 re-solve's per-workgroup areas
    if (link_certifies(n, t->precision)) {
        if (n > 64) {
            L.resolve_blocks = resolve_generic_blocks(resolve_generic_area_bytes(n), area_budget);
            b += (size_t)L.resolve_blocks * resolve_generic_area_bytes(n);
        } else {
            L.resolve_blocks = resolve_small_blocks(n, area_budget);
            b += (size_t)L.resolve_blocks * (n <= 32 ? 2 : 1) * resolve_small_area_bytes(n);
        }
    }
    L.total = b;
    return L;
}

size_t link_workspace_for_chunk(const fringe_link_tile* t, int chunk, size_t area_budget) {
    return link_layout(t, chunk, area_budget).total;
}

// Largest chunk (whole rows, whole tiles of the staged Gram when possible)
// whose workspace fits in `bytes`; 0 when not even one row fits.
int link_chunk_for_bytes(const fringe_link_tile* t, size_t bytes, size_t area_budget) {
    const long long out_rows = t->row_end - t->row_begin;
    if (out_rows <= 0) return t->cols;
    const int th = link_uses_mfma(t->nbands, t->precision) ? gram_stage_rows(t->nbands, t->Nx, t->Ny) : 1;
    long long lo = 0, hi = out_rows;   // lo rows fit (0 = nothing yet)
    while (lo < hi) {
        const long long mid = (lo + hi + 1) / 2;
        if (link_workspace_for_chunk(t, (int)(mid * t->cols), area_budget) <= bytes) lo = mid;
        else hi = mid - 1;
    }
    if (lo > th && lo < out_rows) lo -= lo % th;
    return (int)(lo * t->cols);
}

// Body of fringe_link_run with an explicit chunk (pixels per Gram/solve
// round, a multiple of cols); the workspace must hold link_layout(chunk).
int link_run_chunked(const fringe_link_tile* t, const void* d_slc, const uint32_t* d_wts, void* d_phase,
                     float* d_tcorr, void* d_comp, void* d_work, size_t work_bytes, int chunk, void* stream,
                     size_t area_budget) {
    const LinkLayout LL = link_layout(t, chunk, area_budget);
    if (work_bytes < LL.total) return FRINGE_ERR_WORKSPACE;
    if (t->row_end == t->row_begin) return 0;
    LinkPlan P{};
    P.rows = t->rows;
    P.cols = t->cols;
    P.n = t->nbands;
    P.Nx = t->Nx;
    P.Ny = t->Ny;
    P.nul = shp_words(t->Nx, t->Ny);
    P.module = t->module;
    P.method = t->method;
    // evd.cpp:566 ignores minNeighbors and requires 2; phase_link.cpp:524 honours it
    P.min_nb = (t->module == FRINGE_MODULE_EVD) ? 2 : t->min_neighbors;
    P.ref = t->ref_index;
    P.bw = t->bandwidth;
    P.precision = t->precision;
    P.row_begin = t->row_begin;
    P.row_end = t->row_end;
    P.npx = (long long)t->rows * t->cols;
    P.chunk = chunk;
    P.d_gamma = (char*)d_work + LL.gamma;
    P.d_npix = (int*)((char*)d_work + LL.npix);
    P.d_pack = (uint32_t*)((char*)d_work + LL.pack);
    if (link_certifies(P.n, P.precision)) {
        P.d_cert_count = (int*)((char*)d_work + LL.cert);
        P.d_cert_list = (int*)((char*)d_work + LL.cert + 256);
        P.cert_cap = (int)((long long)(t->row_end - t->row_begin) * t->cols);
    }
    if (LL.awork_stride) {
        P.d_awork = (char*)d_work + LL.awork;
        P.awork_stride = LL.awork_stride;
    }
    if (link_certifies(P.n, P.precision)) {
        P.d_rwork = (char*)d_work + LL.rwork;
        P.resolve_blocks = LL.resolve_blocks;
    }
    P.d_queue = (int*)((char*)d_work + LL.queue);
    hipStream_t s = (hipStream_t)stream;
    if (link_launch(P, (const float2*)d_slc, d_wts, (float2*)d_phase, d_tcorr, (float2*)d_comp, s) != 0) {
        set_last_error("fringe_link_run", hipGetLastError());
        return FRINGE_ERR_DEVICE;
    }
    return 0;
}

}  // namespace fringe

using namespace fringe;

extern "C" {

const char* fringe_version(void) { return "fringe_amd 0.1 (gfx950)"; }
const char* fringe_last_error(void) { return g_last_error.c_str(); }
int fringe_mask_words(int Nx, int Ny) { return shp_words(Nx, Ny); }

size_t fringe_link_workspace_bytes(const fringe_link_tile* t) {
    if (link_check(t) != 0) return 0;
    return link_workspace_for_chunk(t, link_chunk_pixels(t));
}

int fringe_link_run(const fringe_link_tile* t, const void* d_slc, const uint32_t* d_wts,
                    void* d_phase, float* d_tcorr, void* d_comp, void* d_work, size_t work_bytes,
                    void* stream) {
    const int st = link_check(t);
    if (st != 0) return st;
    if (!d_slc || !d_wts || !d_phase || !d_tcorr || !d_comp || !d_work) return FRINGE_ERR_ARGUMENT;
    return link_run_chunked(t, d_slc, d_wts, d_phase, d_tcorr, d_comp, d_work, work_bytes, link_chunk_pixels(t),
                            stream);
}

long long fringe_link_resolved_pixels(const fringe_link_tile* t, const void* d_work, size_t work_bytes) {
    if (link_check(t) != 0 || !d_work) return -1;
    if (!link_certifies(t->nbands, t->precision)) return 0;
    if (work_bytes < 256) return -1;
    int c = 0;   // the count sits at the start of the workspace
    if (hipMemcpy(&c, d_work, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return c;
}

// ---- batch of independent tiles --------------------------------------------
// Layout: queue counters; per tile its certification list and packed
// records; the round buffer (coherence, and work matrices above 64 dates in
// global memory) of up to ~2 GB; the round's neighbour counts; the FP64
// re-solve areas above 64 dates.
struct BatchLayout {
    std::vector<size_t> cert, pack;
    std::vector<size_t> astride;
    size_t queue, round, round_bytes, npix, rwork, total;
};

static int batch_check(const fringe_link_batch_item* items, int count) {
    if (!items || count <= 0) return FRINGE_ERR_ARGUMENT;
    for (int i = 0; i < count; ++i) {
        const int st = link_check(&items[i].tile);
        if (st) return st;
        if (items[i].tile.module != items[0].tile.module || items[i].tile.method != items[0].tile.method ||
            items[i].tile.precision != items[0].tile.precision)
            return FRINGE_ERR_UNSUPPORTED;
    }
    return 0;
}

static BatchLayout batch_layout(const fringe_link_batch_item* items, int count) {
    BatchLayout L;
    size_t b = 0;
    L.queue = b;
    b += 256;
    size_t need = 0, max_row = 0, max_rwork = 0, pixels = 0;
    for (int i = 0; i < count; ++i) {
        const fringe_link_tile* t = &items[i].tile;
        const int n = t->nbands;
        const long long out = (long long)(t->row_end - t->row_begin) * t->cols;
        L.cert.push_back(b);
        if (link_certifies(n, t->precision)) b += 256 + align256((size_t)out * sizeof(int));
        L.pack.push_back(b);
        if (link_uses_mfma(n, t->precision)) b += align256(mfma_pack_bytes(t->rows, t->cols, n));
        size_t as = 0;
        if (n > 64 && !solve_a_in_lds(n, t->precision))
            as = align256((size_t)n * (n | 1) * (t->precision == FRINGE_PREC_FP64 ? 16 : 8));
        L.astride.push_back(as);
        const size_t per_px = link_gamma_bytes_per_pixel(n, t->precision) - sizeof(int) + as;
        need += (size_t)out * per_px + 512;
        max_row = std::max(max_row, (size_t)t->cols * per_px + 1024);
        pixels += (size_t)out;
        if (link_certifies(n, t->precision))
            max_rwork = std::max(max_rwork, n > 64 ? (size_t)resolve_generic_blocks(resolve_generic_area_bytes(n)) *
                                                         resolve_generic_area_bytes(n)
                                                   : (size_t)resolve_small_blocks(n, kResolveAreaBudget) *
                                                         (n <= 32 ? 2 : 1) * resolve_small_area_bytes(n));
    }
    L.round = b;
    L.round_bytes = std::max(std::min(need, (size_t)2048 << 20), max_row);
    b += align256(L.round_bytes);
    L.npix = b;
    b += align256(pixels * sizeof(int));
    L.rwork = b;
    b += max_rwork;
    L.total = b;
    return L;
}

extern "C" size_t fringe_link_batch_workspace_bytes(const fringe_link_batch_item* items, int count) {
    if (batch_check(items, count) != 0) return 0;
    return batch_layout(items, count).total;
}

extern "C" int fringe_link_batch_run(const fringe_link_batch_item* items, int count, void* d_work,
                                     size_t work_bytes, void* stream) {
    const int st = batch_check(items, count);
    if (st) return st;
    const BatchLayout L = batch_layout(items, count);
    if (!d_work || work_bytes < L.total) return FRINGE_ERR_WORKSPACE;
    BatchPlan B;
    B.module = items[0].tile.module;
    B.method = items[0].tile.method;
    B.precision = items[0].tile.precision;
    char* w = (char*)d_work;
    B.d_queue = (int*)(w + L.queue);
    B.d_round = w + L.round;
    B.round_bytes = L.round_bytes;
    B.d_npix = (int*)(w + L.npix);
    B.d_rwork = L.rwork < L.total ? w + L.rwork : nullptr;
    for (int i = 0; i < count; ++i) {
        const fringe_link_batch_item& it = items[i];
        const fringe_link_tile* t = &it.tile;
        if (!it.d_slc || !it.d_wts || !it.d_phase || !it.d_tcorr || !it.d_comp) return FRINGE_ERR_ARGUMENT;
        BatchTile T{};
        LinkPlan& P = T.P;
        P.rows = t->rows;
        P.cols = t->cols;
        P.n = t->nbands;
        P.Nx = t->Nx;
        P.Ny = t->Ny;
        P.nul = shp_words(t->Nx, t->Ny);
        P.module = t->module;
        P.method = t->method;
        P.min_nb = (t->module == FRINGE_MODULE_EVD) ? 2 : t->min_neighbors;
        P.ref = t->ref_index;
        P.bw = t->bandwidth;
        P.precision = t->precision;
        P.row_begin = t->row_begin;
        P.row_end = t->row_end;
        P.npx = (long long)t->rows * t->cols;
        P.d_pack = (uint32_t*)(w + L.pack[i]);
        if (link_certifies(P.n, P.precision)) {
            P.d_cert_count = (int*)(w + L.cert[i]);
            P.d_cert_list = (int*)(w + L.cert[i] + 256);
            P.cert_cap = (int)((long long)(t->row_end - t->row_begin) * t->cols);
        }
        P.awork_stride = L.astride[i];
        T.slc = (const float2*)it.d_slc;
        T.wts = it.d_wts;
        T.phase = (float2*)it.d_phase;
        T.tcorr = it.d_tcorr;
        T.comp = (float2*)it.d_comp;
        B.tiles.push_back(T);
    }
    if (link_batch_launch(B, (hipStream_t)stream) != 0) {
        set_last_error("fringe_link_batch_run", hipGetLastError());
        return FRINGE_ERR_DEVICE;
    }
    return 0;
}

size_t fringe_coherence_workspace_bytes(const fringe_link_tile* t) {
    if (link_check(t) != 0) return 0;
    return link_uses_mfma(t->nbands, t->precision) ? align256(mfma_pack_bytes(t->rows, t->cols, t->nbands)) : 256;
}

int fringe_coherence_run(const fringe_link_tile* t, const void* d_slc, const uint32_t* d_wts, void* d_gamma,
                         int* d_npix, void* d_work, size_t work_bytes, void* stream) {
    const int st = link_check(t);
    if (st != 0) return st;
    if (!d_slc || !d_wts || !d_gamma || !d_npix || !d_work) return FRINGE_ERR_ARGUMENT;
    if (work_bytes < fringe_coherence_workspace_bytes(t)) return FRINGE_ERR_WORKSPACE;
    LinkPlan P{};
    P.rows = t->rows;
    P.cols = t->cols;
    P.n = t->nbands;
    P.Nx = t->Nx;
    P.Ny = t->Ny;
    P.nul = shp_words(t->Nx, t->Ny);
    P.module = t->module;
    P.method = t->method;
    P.min_nb = (t->module == FRINGE_MODULE_EVD) ? 2 : t->min_neighbors;
    P.precision = t->precision;
    P.row_begin = t->row_begin;
    P.row_end = t->row_end;
    P.npx = (long long)t->rows * t->cols;
    P.d_pack = (uint32_t*)d_work;
    if (coherence_launch(P, (const float2*)d_slc, d_wts, d_gamma, d_npix, (hipStream_t)stream) != 0) {
        set_last_error("fringe_coherence_run", hipGetLastError());
        return FRINGE_ERR_DEVICE;
    }
    return 0;
}

double fringe_ks_table(int n, double prob, uint8_t* h_table) {
    const double rn = n;
    const double scale = std::sqrt(rn * rn / (rn + rn));   // KS2sample.hpp:139
    double margin = INFINITY;
    for (int k = 0; k <= n; ++k) {
        const double rd = (double)k / rn;
        const double p = kolmogorov_prob(rd * scale);
        if (h_table) h_table[k] = (p >= prob) ? 1 : 0;     // nmap.cpp:462
        margin = std::min(margin, std::fabs(p - prob));
    }
    return margin;
}

// Workspace of fringe_nmap_run: sorted amplitudes [rows*cols][n] float,
// validity [rows*cols], the AD2 term table, then the forward-half mask
// words and counts.  With P, points its device pointers into d_work.
static size_t nmap_layout(const fringe_nmap_tile* t, NmapPlan* P, void* d_work = nullptr) {
    const size_t npx = (size_t)t->rows * t->cols;
    const size_t o_valid = align256(npx * t->nbands * sizeof(float));
    const size_t o_ad2 = o_valid + align256(npx);
    const size_t o_fwd = o_ad2 + (t->method == 1 ? align256((size_t)(2 * t->nbands - 1) * (t->nbands + 1) * sizeof(double)) : 0);
    const size_t o_fcnt = o_fwd + align256(npx * nmap_fwd_words(t->Nx, t->Ny) * sizeof(uint32_t));
    const size_t end = o_fcnt + align256(npx * sizeof(int32_t));
    if (P) {
        char* w = (char*)d_work;
        P->d_amp = (float*)w;
        P->d_valid = (uint8_t*)(w + o_valid);
        P->d_ad2 = t->method == 1 ? (double*)(w + o_ad2) : nullptr;
        P->d_fwd = (uint32_t*)(w + o_fwd);
        P->d_fcnt = (int32_t*)(w + o_fcnt);
    }
    return end;
}

size_t fringe_nmap_workspace_bytes(const fringe_nmap_tile* t) {
    if (!t || t->rows <= 0 || t->cols <= 0 || t->nbands <= 0) return 0;
    return nmap_layout(t, nullptr);
}

double fringe_ad2_threshold(int n, double prob, double* h_sigma) {
    if (n < 2) return NAN;
    if (h_sigma) *h_sigma = ad2_sigma(n, n);
    return ad2_threshold(prob);
}

int fringe_nmap_run(const fringe_nmap_tile* t, const void* d_slc, const uint8_t* d_mask,
                    const double* h_amp_const, uint32_t* d_wts, int32_t* d_count, void* d_work,
                    size_t work_bytes, void* stream) {
    if (!t || !d_slc || !d_wts || !d_count || !d_work) return FRINGE_ERR_ARGUMENT;
    if (t->rows <= 0 || t->cols <= 0 || t->nbands <= 0 || t->Nx < 0 || t->Ny < 0)
        return FRINGE_ERR_ARGUMENT;
    if (t->row_begin < 0 || t->row_end > t->rows || t->row_begin > t->row_end) return FRINGE_ERR_ARGUMENT;
    if (t->nbands > kMaxBandsNmap) return FRINGE_ERR_UNSUPPORTED;
    if (shp_words(t->Nx, t->Ny) > kMaxMaskWords) return FRINGE_ERR_UNSUPPORTED;
    if (t->method != 0 && t->method != 1) return FRINGE_ERR_ARGUMENT;
    if (t->method == 1 && t->nbands < 4) return FRINGE_ERR_UNSUPPORTED;   // sigma_N needs 2n > 3
    if ((long long)t->rows * t->cols >= (1ll << 31) - 1) return FRINGE_ERR_UNSUPPORTED;
    if (work_bytes < fringe_nmap_workspace_bytes(t)) return FRINGE_ERR_WORKSPACE;
    NmapPlan P{};
    P.rows = t->rows;
    P.cols = t->cols;
    P.n = t->nbands;
    P.Nx = t->Nx;
    P.Ny = t->Ny;
    P.nul = shp_words(t->Nx, t->Ny);
    P.row_begin = t->row_begin;
    P.row_end = t->row_end;
    // alpha_b = amplitudeConstant_b / amplitudeConstant_1, missing or <= 0
    // read as 1, alpha_1 forced to 1 (nmap.cpp:204-233)
    for (int b = 0; b < t->nbands; ++b) {
        const double c = h_amp_const ? h_amp_const[b] : 0.0;
        P.alpha[b] = (c <= 0.0) ? 1.0 : c;
    }
    const double norm = P.alpha[0];
    for (int b = 0; b < t->nbands; ++b) P.alpha[b] /= norm;
    P.alpha[0] = 1.0;
    uint8_t tab[kMaxBandsNmap + 1];
    fringe_ks_table(t->nbands, t->prob, tab);
    std::memset(P.table.bits, 0, sizeof(P.table.bits));
    for (int k = 0; k <= t->nbands; ++k)
        if (tab[k]) P.table.bits[k >> 5] |= (1u << (k & 31));
    nmap_layout(t, &P, d_work);
    P.method = t->method;
    if (t->method == 1) {
        P.ad2_sigma = ad2_sigma(t->nbands, t->nbands);
        P.ad2_max = ad2_threshold(t->prob);
    }
    if (nmap_launch(P, (const float2*)d_slc, d_mask, d_wts, d_count, (hipStream_t)stream) != 0) {
        set_last_error("fringe_nmap_run", hipGetLastError());
        return FRINGE_ERR_DEVICE;
    }
    return 0;
}

}  // extern "C"
