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
#include <algorithm>

#include "device_common.hpp"
#include "engine_internal.hpp"
#include "link_solve.hpp"
#include "solve32.hpp"
#include "solve128.hpp"
#include "timing.hpp"

namespace fringe {

// ---------------------------------------------------------------------------
// FP64 Gram over a chunk of output pixels, one group of G lanes per pixel
// (the per-pixel body is gram_f64_pixel in gram64.hpp).
// ---------------------------------------------------------------------------
template <int G, int R, int CB>
__global__ __launch_bounds__(64) void gram_f64(GramArgs a) {
    constexpr int GPW = 64 / G;
    const int grp = threadIdx.x / G, g = threadIdx.x % G;
    const int slot = blockIdx.x * GPW + grp;
    extern __shared__ __align__(16) char smem[];
    const bool inslot = slot < a.slots;
    const long long pix = a.first + (inslot ? slot : 0);
    cx<double>* gam = (cx<double>*)a.gamma + (size_t)(inslot ? slot : 0) * a.n * a.n;
    const int np = gram_f64_pixel<G, R, CB>(a, inslot, pix, gam, smem + grp * a.group_bytes);
    if (inslot && g == 0) a.npix[slot] = np;
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
// Above 64 dates: one pixel per two-wave workgroup, rows g + 128 q; the
// work matrix in LDS, or at a.awork when it does not fit there.
template <typename T, typename Tg, int R>
static void launch_solve_generic(SolveArgs a, hipStream_t s) {
    a.group_bytes = solve_group_bytes<T>(a.n, a.awork == nullptr);
    hipLaunchKernelGGL((solve_kernel<T, Tg, 128, R, false>), dim3((unsigned)a.slots), dim3(128), a.group_bytes, s,
                       a);
}

// 32 < n <= 128: the register-row solver (FP32 with two threads per row,
// four up to 48 dates, in the mixed precisions; FP64 with four in the plain
// path and the re-solve) where the schedule FRINGE_SCHED_GENERIC_A (or its
// default) selects it;
// otherwise the one-row-per-lane solver up to 64 dates and a work matrix
// above.
bool rows_solver_selected(int n, int precision) {
    if (n <= 32 || n > kMaxBandsMfma) return false;
    const int pk = precision == FRINGE_PREC_FP64 ? 0 : 1;
    return schedule_lookup(FRINGE_SCHED_GENERIC_A, n, pk, 0, generic_a_default(n, pk)) == 2;
}

static void launch_rows_solver(const SolveArgs& a, bool fp64, hipStream_t s) {
    if (fp64) launch_rows_solver_f64(a, s);
    else launch_rows_solver_f32(a, s);
}

template <typename T, typename Tg>
static void dispatch_generic(const SolveArgs& a, hipStream_t s) {
    if constexpr (sizeof(T) == sizeof(Tg)) {   // FP32 on the MFMA Gram's coherence, or FP64
        if (!a.awork && rows_solver_selected(a.n, sizeof(T) == 8 ? FRINGE_PREC_FP64 : FRINGE_PREC_MIXED)) {
            launch_rows_solver(a, sizeof(T) == 8, s);
            return;
        }
    }
    if (a.n <= 128) launch_solve_generic<T, Tg, 1>(a, s);
    else launch_solve_generic<T, Tg, 2>(a, s);
}

// Static LDS of the solver kernels that run above 64 dates (their reduction
// scratch); the launch checks add it to the dynamic LDS of a pixel.
size_t solve_static_lds(int precision) {
    hipFuncAttributes fa{};
    size_t most = 0;
    const void* ks[2] = {precision == FRINGE_PREC_FP64 ? (const void*)solve_kernel<double, double, 128, 2, false>
                                                        : (const void*)solve_kernel<float, double, 128, 2, false>,
                         precision == FRINGE_PREC_FP64 ? (const void*)solve_kernel<double, double, 128, 1, false>
                                                        : (const void*)solve_kernel<float, float, 128, 1, false>};
    for (const void* k : ks) most = std::max<size_t>(most, hipFuncGetAttributes(&fa, k) == hipSuccess
                                                               ? fa.sharedSizeBytes : 256);
    return most;
}

// Whether the work matrix of the solver above 64 dates fits in LDS.
// The schedule FRINGE_SCHED_GENERIC_A may move it to global memory even
// when it fits: the pixels in flight are then limited by registers, not LDS.
bool solve_a_in_lds(int n, int precision) {
    if (rows_solver_selected(n, precision)) return true;   // no work matrix at all
    if (schedule_lookup(FRINGE_SCHED_GENERIC_A, n, precision == FRINGE_PREC_FP64 ? 0 : 1, 0, 0) == 1) return false;
    const size_t b = precision == FRINGE_PREC_FP64 ? solve_group_bytes<double>(n) : solve_group_bytes<float>(n);
    return b + solve_static_lds(precision) <= (size_t)device_lds_limit();
}

// Per-workgroup global area of the FP64 re-solve above 64 dates: the FP64
// coherence, then the work matrix.
size_t resolve_generic_area_bytes(int n) {
    const int S = n | 1;
    return align256((size_t)n * n * sizeof(cx<double>)) + align256((size_t)n * S * sizeof(cx<double>));
}

size_t resolve_generic_lds_bytes(int n, int Nx, int Ny) {
    return solve_group_bytes<double>(n, false) + 2 * gram_f64_group_bytes(n, Nx, Ny, kResolveCB) +
           solve_static_lds(FRINGE_PREC_FP64);
}

// Two pixels per wave (n <= 32): the tridiagonal reductions process NC
// columns, the fewest of 16, 24, 28, 30, 32 that hold n (the same results).
template <typename T, typename Tg, int NC>
static void launch_reg32(const SolveArgs& a, hipStream_t s) {
    const unsigned blocks = (unsigned)((a.slots + 1) / 2);
    if (a.module == FRINGE_MODULE_PHASE_LINK)
        hipLaunchKernelGGL((solve_reg_kernel<T, Tg, MODE_EMI, 32, false, NC>), dim3(blocks), dim3(64), 0, s, a);
    else if (a.method == FRINGE_METHOD_MLE)
        hipLaunchKernelGGL((solve_reg_kernel<T, Tg, MODE_MLE, 32, false, NC>), dim3(blocks), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL((solve_reg_kernel<T, Tg, MODE_EVD, 32, false, NC>), dim3(blocks), dim3(64), 0, s, a);
}

// n <= 64: register-resident solver, two pixels per wave (n <= 32) or one
template <typename T, typename Tg>
static void dispatch_reg(const SolveArgs& a, hipStream_t s) {
    if (a.n <= 16) {
        launch_reg32<T, Tg, 16>(a, s);
    } else if (a.n <= 24) {
        launch_reg32<T, Tg, 24>(a, s);
    } else if (a.n <= 28) {
        launch_reg32<T, Tg, 28>(a, s);
    } else if (a.n <= 30) {
        launch_reg32<T, Tg, 30>(a, s);
    } else if (a.n <= 32) {
        launch_reg32<T, Tg, 32>(a, s);
    } else {
        const unsigned blocks = (unsigned)a.slots;
        if (a.module == FRINGE_MODULE_PHASE_LINK)
            hipLaunchKernelGGL((solve_reg_kernel<T, Tg, MODE_EMI, 64, false>), dim3(blocks), dim3(64), 0, s, a);
        else if (a.method == FRINGE_METHOD_MLE)
            hipLaunchKernelGGL((solve_reg_kernel<T, Tg, MODE_MLE, 64, false>), dim3(blocks), dim3(64), 0, s, a);
        else
            hipLaunchKernelGGL((solve_reg_kernel<T, Tg, MODE_EVD, 64, false>), dim3(blocks), dim3(64), 0, s, a);
    }
}

template <int NC>
static void launch_resolve32(const SolveArgs& a, dim3 grid, size_t lds, hipStream_t s) {
    if (a.module == FRINGE_MODULE_PHASE_LINK)
        hipLaunchKernelGGL((solve_reg_kernel<double, double, MODE_EMI, 32, true, NC>), grid, dim3(64), lds, s, a);
    else if (a.method == FRINGE_METHOD_MLE)
        hipLaunchKernelGGL((solve_reg_kernel<double, double, MODE_MLE, 32, true, NC>), grid, dim3(64), lds, s, a);
    else
        hipLaunchKernelGGL((solve_reg_kernel<double, double, MODE_EVD, 32, true, NC>), grid, dim3(64), lds, s, a);
}

// FP64 re-solve (Gram and estimator chain of the plain FP64 path) of the
// pixels the FP32 solver of this tile listed as not certified (n <= 64).
// Each lane group's coherence goes to its own global area at P.d_rwork (in
// LDS it allowed two workgroups per CU); the workgroups loop over the list.
static void launch_resolve(const LinkPlan& P, const uint32_t* wts, SolveArgs a, hipStream_t s) {
    GramArgs g{};
    g.slc = a.slc;
    g.wts = wts;
    g.rows = P.rows;
    g.cols = P.cols;
    g.n = P.n;
    g.Nx = P.Nx;
    g.Ny = P.Ny;
    g.nul = P.nul;
    g.min_nb = P.min_nb;
    g.npx = P.npx;
    a.g64 = g;
    a.list = P.d_cert_list;
    a.list_count = P.d_cert_count;
    a.list_cap = P.cert_cap;
    a.cert_list = nullptr;
    a.cert_count = nullptr;
    a.cert_cap = 0;
    a.group_bytes = resolve_group_bytes(P.n, P.Nx, P.Ny);
    a.awork = P.d_rwork;
    a.awork_stride = resolve_small_area_bytes(P.n);
    const int gpw = P.n <= 32 ? 2 : 1;
    const size_t lds = a.group_bytes * gpw;
    const dim3 grid((unsigned)(P.resolve_blocks ? P.resolve_blocks : resolve_small_blocks(P.n, kResolveAreaBudget)));
    if (P.n <= 16) {
        launch_resolve32<16>(a, grid, lds, s);
    } else if (P.n <= 24) {
        launch_resolve32<24>(a, grid, lds, s);
    } else if (P.n <= 28) {
        launch_resolve32<28>(a, grid, lds, s);
    } else if (P.n <= 30) {
        launch_resolve32<30>(a, grid, lds, s);
    } else if (P.n <= 32) {
        launch_resolve32<32>(a, grid, lds, s);
    } else {
        if (a.module == FRINGE_MODULE_PHASE_LINK)
            hipLaunchKernelGGL((solve_reg_kernel<double, double, MODE_EMI, 64, true>), grid, dim3(64), lds, s, a);
        else if (a.method == FRINGE_METHOD_MLE)
            hipLaunchKernelGGL((solve_reg_kernel<double, double, MODE_MLE, 64, true>), grid, dim3(64), lds, s, a);
        else
            hipLaunchKernelGGL((solve_reg_kernel<double, double, MODE_EVD, 64, true>), grid, dim3(64), lds, s, a);
    }
}

// FP64 re-solve above 64 dates: each workgroup loops over its share of the
// list, with its coherence and work matrix in its global area.
static void launch_resolve_generic(const LinkPlan& P, const uint32_t* wts, SolveArgs a, hipStream_t s) {
    GramArgs g{};
    g.slc = a.slc;
    g.wts = wts;
    g.rows = P.rows;
    g.cols = P.cols;
    g.n = P.n;
    g.Nx = P.Nx;
    g.Ny = P.Ny;
    g.nul = P.nul;
    g.min_nb = P.min_nb;
    g.npx = P.npx;
    a.g64 = g;
    a.list = P.d_cert_list;
    a.list_count = P.d_cert_count;
    a.list_cap = P.cert_cap;
    a.cert_list = nullptr;
    a.cert_count = nullptr;
    a.cert_cap = 0;
    a.awork = P.d_rwork;
    a.awork_stride = P.n <= 64 ? resolve_small_area_bytes(P.n) : resolve_generic_area_bytes(P.n);
    a.group_bytes = solve_group_bytes<double>(P.n, false);
    const size_t lds = a.group_bytes + 2 * gram_f64_group_bytes(P.n, P.Nx, P.Ny, kResolveCB);
    const unsigned blocks = (unsigned)(P.resolve_blocks ? P.resolve_blocks
                                       : P.n <= 64 ? resolve_small_blocks(P.n, kResolveAreaBudget)
                                                   : resolve_generic_blocks(a.awork_stride));
    const dim3 grid(blocks);
    if (rows_solver_selected(P.n, FRINGE_PREC_FP64)) {   // the plain FP64 path's solver
        launch_resolve_rows(a, blocks, s);
        return;
    }
    if (P.n <= 128)
        hipLaunchKernelGGL((solve_kernel<double, double, 128, 1, true>), grid, dim3(128), lds, s, a);
    else
        hipLaunchKernelGGL((solve_kernel<double, double, 128, 2, true>), grid, dim3(128), lds, s, a);
}

// LDS of one lane group of the re-solve up to 64 dates: the Gram's scratch.
size_t resolve_group_bytes(int n, int Nx, int Ny) { return gram_f64_group_bytes(n, Nx, Ny, kResolveSmallCB); }

size_t resolve_small_area_bytes(int n) { return align256((size_t)n * n * sizeof(cx<double>)); }

int resolve_small_blocks(int n, size_t budget) {
    const int gpw = n <= 32 ? 2 : 1;
    const size_t b = (budget / 4) / ((size_t)gpw * resolve_small_area_bytes(n));
    return (int)(b < 16 ? 16 : (b > 1024 ? 1024 : b));
}

size_t resolve_lds_bytes(int n, int Nx, int Ny) {
    const int gpw = n <= 32 ? 2 : 1;
    const size_t stat = n <= 32 ? sizeof(LdsR<double, 32>) * 2 : sizeof(LdsR<double, 64>);
    return resolve_group_bytes(n, Nx, Ny) * gpw + stat;
}

__global__ void zero_i32(int* p, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

bool link_uses_mfma(int n, int precision) { return precision != FRINGE_PREC_FP64 && n <= kMaxBandsMfma; }
bool link_certifies(int n, int precision) { return precision != FRINGE_PREC_FP64; }

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
    else if (P.n <= 128) launch_gram_f64<64, 2, 16>(g, s);
    else launch_gram_f64<64, 4, 16>(g, s);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------------------
// Pieces of a link: the Gram of a segment of whole rows, the solver launch
// (a grid, or a persistent queue of jobs), the FP64 re-solve of a tile.
// ---------------------------------------------------------------------------
static void gram_segment(const LinkPlan& P, const float2* slc, const uint32_t* wts, long long first, int slots,
                         void* gamma, int* npix, hipStream_t s, int* rc) {
    if (!link_uses_mfma(P.n, P.precision)) {
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
        else if (P.n <= 128) launch_gram_f64<64, 2, 16>(g, s);
        else launch_gram_f64<64, 4, 16>(g, s);
    } else {
        LinkPlan Q = P;
        Q.d_gamma = gamma;
        Q.d_npix = npix;
        if (gram_mfma_launch(Q, P.d_pack, wts, first, slots, s) != 0) *rc = -1;
    }
}

static int cu_count() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cached[dev]) {
        hipDeviceProp_t prop;
        cached[dev] = (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
                          ? prop.multiProcessorCount : 256;
    }
    return cached[dev];
}

// Date class of the solvers: 0 register NR = 32 (n <= 32), 1 register
// NR = 64 (n <= 64), 2 generic one row per thread (n <= 128: the MFMA
// Gram's complex64 coherence in the mixed precisions), 3 generic two rows
// per thread (n > 128, complex128 coherence).
static int solve_class(int n) { return n <= 32 ? 0 : (n <= 64 ? 1 : (n <= kMaxBandsMfma ? 2 : 3)); }

static PJob make_job(const LinkPlan& P, const float2* slc, float2* phase, float* tcorr, float2* comp,
                     long long first, int slots, const void* gamma, const int* npix, void* awork, size_t awork_stride) {
    PJob J{};
    J.gamma = gamma;
    J.npix = npix;
    J.slc = slc;
    J.phase = phase;
    J.tcorr = tcorr;
    J.comp = comp;
    J.cert_list = P.d_cert_list;
    J.cert_count = P.d_cert_count;
    J.cert_cap = P.cert_cap;
    J.awork = awork;
    J.awork_stride = awork_stride;
    J.npx = P.npx;
    J.first = first;
    J.slots = slots;
    J.n = P.n;
    J.ref = P.ref;
    J.bw = P.bw;
    const bool f64 = P.precision == FRINGE_PREC_FP64;
    J.group_bytes = (unsigned)(f64 ? solve_group_bytes<double>(P.n, awork == nullptr)
                                   : solve_group_bytes<float>(P.n, awork == nullptr));
    return J;
}

template <typename K> static int occupancy_blocks(K kernel, int threads, size_t lds) {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kernel, threads, lds) != hipSuccess || b < 1) b = 1;
    return b;
}

// One persistent launch over the jobs of pa (all of class cls).  sched: 0
// = default (one full residency of the kernel, grab 4), else
// (waves_per_cu << 8 | grab).
template <typename T, typename Tg>
static void launch_persistent(PersistArgs pa, int cls, int mode, int sched, hipStream_t s) {
    int grab = 4, waves_cu = 0;
    if (sched) {
        waves_cu = sched >> 8;
        grab = sched & 0xff;
    }
    pa.grab = grab;
    int units = 0;
    size_t lds = 0;
    for (int j = 0; j < pa.njobs; ++j) {
        const int u = cls == 0 ? (pa.job[j].slots + 1) / 2 : pa.job[j].slots;
        units += u;
        pa.unit_end[j] = units;
        lds = std::max<size_t>(lds, pa.job[j].group_bytes);
    }
    if (units == 0) return;
    hipLaunchKernelGGL(zero_i32, dim3(1), dim3(64), 0, s, pa.queue, 1);
    const int cus = cu_count();
    auto grid_for = [&](auto kernel, int threads, size_t dyn) {
        const int per_cu = waves_cu ? std::max(1, waves_cu / (threads / 64)) : occupancy_blocks(kernel, threads, dyn);
        return dim3((unsigned)std::min<long long>((long long)per_cu * cus, (units + grab - 1) / grab));
    };
    if (cls == 2) {   // complex64 coherence in the mixed precisions, complex128 in FP64
        if constexpr (sizeof(T) == sizeof(Tg)) {
            auto k = solve_generic_persistent<T, Tg, 1>;
            hipLaunchKernelGGL(k, grid_for(k, 128, lds), dim3(128), lds, s, pa);
        }
        return;
    }
    if (cls == 3) {   // complex128 coherence (FP64 Gram) in every precision
        if constexpr (sizeof(Tg) == 8) {
            auto k = solve_generic_persistent<T, Tg, 2>;
            hipLaunchKernelGGL(k, grid_for(k, 128, lds), dim3(128), lds, s, pa);
        }
        return;
    }
    if constexpr (sizeof(T) == sizeof(Tg)) {
#define FRINGE_PERSIST(MODE_, NR_)                                                   \
    do {                                                                             \
        auto k = solve_reg_persistent<T, Tg, MODE_, NR_>;                            \
        hipLaunchKernelGGL(k, grid_for(k, 64, 0), dim3(64), 0, s, pa);              \
    } while (0)
        if (cls == 0) {
            if (mode == 0) FRINGE_PERSIST(MODE_EMI, 32);
            else if (mode == 2) FRINGE_PERSIST(MODE_MLE, 32);
            else FRINGE_PERSIST(MODE_EVD, 32);
        } else {
            if (mode == 0) FRINGE_PERSIST(MODE_EMI, 64);
            else if (mode == 2) FRINGE_PERSIST(MODE_MLE, 64);
            else FRINGE_PERSIST(MODE_EVD, 64);
        }
#undef FRINGE_PERSIST
    }
}

static void launch_persistent_any(const PersistArgs& pa, int cls, int mode, int precision, int sched, hipStream_t s) {
    if (precision == FRINGE_PREC_FP64) launch_persistent<double, double>(pa, cls, mode, sched, s);
    else if (cls == 3) launch_persistent<float, double>(pa, cls, mode, sched, s);
    else launch_persistent<float, float>(pa, cls, mode, sched, s);
}

static void resolve_tile(const LinkPlan& P, const float2* slc, const uint32_t* wts, float2* phase, float* tcorr,
                         float2* comp, hipStream_t s) {
    StageScope sc(STAGE_RESOLVE, s);
    SolveArgs a{};
    a.slc = slc;
    a.npx = P.npx;
    a.n = P.n;
    a.module = P.module;
    a.method = P.method;
    a.ref = P.ref;
    a.bw = P.bw;
    a.phase = phase;
    a.tcorr = tcorr;
    a.comp = comp;
    if (P.n <= 64 && !rows_solver_selected(P.n, FRINGE_PREC_FP64)) launch_resolve(P, wts, a, s);
    else launch_resolve_generic(P, wts, a, s);
}

int link_launch(const LinkPlan& P, const float2* slc, const uint32_t* wts, float2* phase,
                float* tcorr, float2* comp, hipStream_t s) {
    const long long first_out = (long long)P.row_begin * P.cols;
    const long long total = (long long)(P.row_end - P.row_begin) * P.cols;
    const bool fp64 = (P.precision == FRINGE_PREC_FP64);
    const bool mfma = link_uses_mfma(P.n, P.precision);
    const int mode = solve_mode_index(P.module, P.method);
    const int sched = schedule_lookup(FRINGE_SCHED_SOLVE, P.n, mode, fp64 ? 0 : 1, 0);
    if (P.d_cert_count) {
        hipLaunchKernelGGL(zero_i32, dim3(1), dim3(64), 0, s, P.d_cert_count, 1);
    }
    if (mfma) {
        StageScope sc(STAGE_PACK, s);
        if (pack_launch(slc, P.rows, P.cols, P.n, P.precision != FRINGE_PREC_MIXED_BF16, P.d_pack, s) != 0)
            return -1;
    }
    int rc = 0;
    for (long long done = 0; done < total; done += P.chunk) {
        const int slots = (int)((total - done) < P.chunk ? (total - done) : P.chunk);
        const long long first = first_out + done;
        {
            StageScope sc(STAGE_GRAM, s);
            gram_segment(P, slc, wts, first, slots, P.d_gamma, P.d_npix, s, &rc);
            if (rc) return -1;
        }
        StageScope sc(STAGE_SOLVE, s);
        if (sched && P.d_queue && !rows_solver_selected(P.n, P.precision)) {
            PersistArgs pa{};
            pa.module = P.module;
            pa.method = P.method;
            pa.njobs = 1;
            pa.queue = P.d_queue;
            pa.job[0] = make_job(P, slc, phase, tcorr, comp, first, slots, P.d_gamma, P.d_npix, P.d_awork,
                                 P.awork_stride);
            launch_persistent_any(pa, solve_class(P.n), mode, P.precision, sched, s);
        } else {
            SolveArgs a{};
            a.cert_list = P.d_cert_list;
            a.cert_count = P.d_cert_count;
            a.cert_cap = P.cert_cap;
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
            if (P.n <= 64 && !rows_solver_selected(P.n, P.precision)) {
                if (fp64) dispatch_reg<double, double>(a, s);
                else dispatch_reg<float, float>(a, s);
            } else {
                a.awork = P.d_awork;
                a.awork_stride = P.awork_stride;
                if (fp64) dispatch_generic<double, double>(a, s);
                else if (mfma) dispatch_generic<float, float>(a, s);
                else dispatch_generic<float, double>(a, s);
            }
        }
        if (hipGetLastError() != hipSuccess) return -1;
    }
    if (P.d_cert_count) {
        resolve_tile(P, slc, wts, phase, tcorr, comp, s);
        if (hipGetLastError() != hipSuccess) return -1;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Batch of independent tiles (fringe_link_batch_run): rounds of coherence
// segments from several tiles in one buffer, one persistent solve launch
// per date class and round.
// ---------------------------------------------------------------------------
static SolveArgs solve_args_of(const PJob& J, int module, int method) {
    SolveArgs a{};
    a.gamma = J.gamma;
    a.npix = J.npix;
    a.slc = J.slc;
    a.npx = J.npx;
    a.first = J.first;
    a.slots = J.slots;
    a.n = J.n;
    a.module = module;
    a.method = method;
    a.ref = J.ref;
    a.bw = J.bw;
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

int link_batch_launch(const BatchPlan& B, hipStream_t s) {
    const int count = (int)B.tiles.size();
    for (int i = 0; i < count; ++i) {
        const BatchTile& T = B.tiles[i];
        if (T.P.d_cert_count) hipLaunchKernelGGL(zero_i32, dim3(1), dim3(64), 0, s, T.P.d_cert_count, 1);
        if (link_uses_mfma(T.P.n, T.P.precision)) {
            StageScope sc(STAGE_PACK, s);
            if (pack_launch(T.slc, T.P.rows, T.P.cols, T.P.n, T.P.precision != FRINGE_PREC_MIXED_BF16, T.P.d_pack,
                            s) != 0)
                return -1;
        }
    }
    const int mode = solve_mode_index(B.module, B.method);
    int item = 0;
    long long row = count ? B.tiles[0].P.row_begin : 0;
    int rc = 0;
    while (item < count) {
        // classes 0..3 by date count (solve_class); 4: the register-row
        // solver's tiles, one grid launch each
        PersistArgs pa[5] = {};
        size_t off = 0;      // bytes of the round buffer in use
        long long soff = 0;  // slots in use
        bool full = false;
        while (item < count && !full) {
            const BatchTile& T = B.tiles[item];
            const LinkPlan& P = T.P;
            const int cls = rows_solver_selected(P.n, P.precision) ? 4 : solve_class(P.n);
            const long long rem = P.row_end - row;
            if (rem <= 0) {
                ++item;
                if (item < count) row = B.tiles[item].P.row_begin;
                continue;
            }
            if (pa[cls].njobs == kMaxJobs) break;
            // bytes per row: coherence matrices, plus the work matrices above 64
            // dates when they live in global memory, each block 256-aligned
            const size_t per_row = (size_t)P.cols * (link_gamma_bytes_per_pixel(P.n, P.precision) - sizeof(int) +
                                                     P.awork_stride);
            long long rows = (long long)((B.round_bytes > off + 512 ? B.round_bytes - off - 512 : 0) / per_row);
            rows = std::min(rows, rem);
            const int th = link_uses_mfma(P.n, P.precision) ? gram_stage_rows(P.n, P.Nx, P.Ny) : 1;
            if (rows < rem && rows > th) rows -= rows % th;
            if (rows <= 0) {
                full = true;
                break;
            }
            const int slots = (int)(rows * P.cols);
            const long long first = row * P.cols;
            char* gam = (char*)B.d_round + off;
            const size_t gbytes = align256((size_t)slots * P.n * P.n * (link_uses_mfma(P.n, P.precision) ? 8 : 16));
            void* aw = nullptr;
            if (P.awork_stride) aw = gam + gbytes;
            int* npix = B.d_npix + soff;
            {
                StageScope sc(STAGE_GRAM, s);
                gram_segment(P, T.slc, T.wts, first, slots, gam, npix, s, &rc);
                if (rc) return -1;
            }
            pa[cls].job[pa[cls].njobs++] = make_job(P, T.slc, T.phase, T.tcorr, T.comp, first, slots, gam, npix, aw,
                                                    P.awork_stride);
            off += gbytes + (size_t)slots * P.awork_stride;
            soff += slots;
            row += rows;
        }
        StageScope sc(STAGE_SOLVE, s);
        for (int j = 0; j < pa[4].njobs; ++j)
            launch_rows_solver(solve_args_of(pa[4].job[j], B.module, B.method), B.precision == FRINGE_PREC_FP64, s);
        for (int c = 0; c < 4; ++c) {
            if (!pa[c].njobs) continue;
            pa[c].module = B.module;
            pa[c].method = B.method;
            pa[c].queue = B.d_queue + c;
            const int sched = schedule_lookup(FRINGE_SCHED_SOLVE, c == 0 ? 32 : (c == 1 ? 64 : 256), mode,
                                              B.precision == FRINGE_PREC_FP64 ? 0 : 1, 0);
            if (pa[c].njobs == 1 && sched == 0) {
                // a lone job of its class: the launch the schedule prefers for one tile
                const SolveArgs a = solve_args_of(pa[c].job[0], B.module, B.method);
                const bool f64 = B.precision == FRINGE_PREC_FP64;
                if (c < 2) {
                    if (f64) dispatch_reg<double, double>(a, s);
                    else dispatch_reg<float, float>(a, s);
                } else if (f64) {
                    dispatch_generic<double, double>(a, s);
                } else if (c == 2) {   // the MFMA Gram's complex64 coherence
                    dispatch_generic<float, float>(a, s);
                } else {
                    dispatch_generic<float, double>(a, s);
                }
            } else {
                launch_persistent_any(pa[c], c, mode, B.precision, sched, s);
            }
        }
        if (hipGetLastError() != hipSuccess) return -1;
    }
    for (int i = 0; i < count; ++i) {
        const BatchTile& T = B.tiles[i];
        if (T.P.d_cert_count && T.P.row_end > T.P.row_begin) {
            LinkPlan Q = T.P;
            Q.d_rwork = B.d_rwork;
            resolve_tile(Q, T.slc, T.wts, T.phase, T.tcorr, T.comp, s);
            if (hipGetLastError() != hipSuccess) return -1;
        }
    }
    return 0;
}

}  // namespace fringe
