// engine_internal.hpp -- plans shared between the host engine and the
// kernel launchers (not part of the public C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "fringe_amd.h"

namespace fringe {

constexpr int kMaxBandsNmap = 128;   // nmap sorts at most this many dates
constexpr int kMaxBandsLink = 256;   // phase-linking solver limit
constexpr int kMaxBandsMfma = 128;   // MFMA Gram (mixed precisions); the FP64 Gram above
// Workgroups of the FP64 re-solve above 64 dates: as many as `budget`
// bytes of per-workgroup global areas allow, between 256 and 1024 (several
// per CU: the FP64 solver with its work matrix in global memory is latency
// bound).  fringe_link_run's budget is 512 MB; the fixed-scratch entry
// point passes a smaller one.
constexpr size_t kResolveAreaBudget = (size_t)512 << 20;
inline int resolve_generic_blocks(size_t area_bytes, size_t budget = kResolveAreaBudget) {
    const size_t b = budget / (area_bytes ? area_bytes : 1);
    return (int)(b < 256 ? 256 : (b > 1024 ? 1024 : b));
}
constexpr int kMaxMaskWords = 64;    // window up to 2048 positions (the reference sequential default 59 x 19 is 1121)

// Decision table of the integer KS statistic, one bit per k = D*n.
struct KsTable {
    uint32_t bits[(kMaxBandsNmap + 1 + 31) / 32];
};

struct NmapPlan {
    int rows, cols, n, Nx, Ny, nul, row_begin, row_end;
    double alpha[kMaxBandsNmap];  // amplitudeConstant_b / amplitudeConstant_1
    KsTable table;
    float* d_amp;      // [rows*cols][n] sorted amplitudes
    uint8_t* d_valid;  // [rows*cols]
    int method;        // 0 = KS2, 1 = AD2
    double ad2_sigma;  // AD2: standard deviation of A2_kN (k = 2, n + n samples)
    double ad2_max;    // AD2: similar iff the standardised statistic <= ad2_max
    double* d_ad2;     // AD2: [2n-1][n+1] per-position terms of the inner sum
    uint32_t* d_fwd;   // forward-half mask words, [rows*cols][nul - centre/32]
    int32_t* d_fcnt;   // forward-half counts, [rows*cols]
};

// uint32 words per pixel of the forward half of a mask (the centre's word on)
int nmap_fwd_words(int Nx, int Ny);

int nmap_launch(const NmapPlan& P, const float2* slc, const uint8_t* mask, uint32_t* wts,
                int32_t* count, hipStream_t s);

struct LinkPlan {
    int rows, cols, n, Nx, Ny, nul;
    int module, method, min_nb, ref, bw, precision;
    int row_begin, row_end;
    long long npx;
    int chunk;          // pixels per Gram/solve round
    void* d_gamma;      // [chunk][n][n] complex (double for FP64, float for mixed)
    int* d_npix;        // [chunk]
    uint32_t* d_pack;   // mixed precision: packed (hi|lo) records, [rows*cols+1][2n]
    // mixed precision: pixels whose FP32 result is not certified, re-solved
    // in FP64 at the end of the tile (null when not used)
    int* d_cert_list;   // [cert_cap]
    int* d_cert_count;  // [1]
    int cert_cap;       // output pixels of the tile
    // above 64 dates: the solver's work matrices when they do not fit in LDS
    // ([chunk] areas of awork_stride bytes, null otherwise), and the FP64
    // re-solve's per-workgroup areas (mixed precisions)
    void* d_awork;
    size_t awork_stride;
    void* d_rwork;
    int resolve_blocks;  // workgroups (and areas at d_rwork) of that re-solve; 0 = resolve_generic_blocks
    int* d_queue;       // work queue counter of a persistent solve launch
};

// A batch of independent tiles (fringe_link_batch_run): per tile its plan
// (pack records, certification list, FP64 re-solve areas) and buffers; the
// rounds share one coherence buffer and one neighbour-count array.
struct BatchTile {
    LinkPlan P;
    const float2* slc;
    const uint32_t* wts;
    float2* phase;
    float* tcorr;
    float2* comp;
};
struct BatchPlan {
    std::vector<BatchTile> tiles;
    int module, method, precision;
    void* d_round;       // round buffer: coherence (and work) matrices of the segments
    size_t round_bytes;
    int* d_npix;         // neighbour counts of a round's slots
    int* d_queue;        // [3] queue counters, one per date class
    void* d_rwork;       // FP64 re-solve areas above 64 dates
};
int link_batch_launch(const BatchPlan& B, hipStream_t s);

// The FP64 re-solve of uncertified pixels (n <= 64): dynamic LDS per lane
// group (its coherence plus the FP64 Gram's scratch), and the whole
// workgroup's LDS including the solver's static part.
size_t resolve_group_bytes(int n, int Nx, int Ny);
// Re-solve up to 64 dates: one lane group's coherence area in the workspace,
// and the workgroups that a quarter of `budget` of such areas allows
// (16..1024; two groups per workgroup up to 32 dates).
size_t resolve_small_area_bytes(int n);
int resolve_small_blocks(int n, size_t budget);
// Column block of that re-solve's FP64 Gram: each coherence entry sums the
// same products in the same order whatever the block (bit-identical to the
// plain path's 32); 8 keeps the accumulators, and the kernel, within the
// plain FP64 solver's registers.
constexpr int kResolveSmallCB = 8;
size_t resolve_lds_bytes(int n, int Nx, int Ny);
// Whether a tile of this precision and date count certifies its FP32 results.
bool link_certifies(int n, int precision);
int device_lds_limit();
// Above 64 dates: whether the solver's work matrix fits in LDS; the FP64
// re-solve's global area per workgroup and its LDS.
bool solve_a_in_lds(int n, int precision);
size_t resolve_generic_area_bytes(int n);
size_t resolve_generic_lds_bytes(int n, int Nx, int Ny);
inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
// Value of a schedule key (fringe_schedule_set), or fallback when unset.
int schedule_lookup(int kind, int k0, int k1, int k2, int fallback);
// FRINGE_SCHED_GENERIC_A when unset: the register-row solver (2) at
// 33..128 dates (FP32: 33..48 and 65..128), otherwise the one-row-per-lane
// solver up to 64 dates and the LDS work matrix above (0).
constexpr int kGenericARowsDefault = 2;
int generic_a_default(int n, int precision_key);
// 33..128 dates under FRINGE_SCHED_GENERIC_A = 2 (either precision; the
// FP64 re-solve follows the FP64 choice, whose outputs it must equal).
bool rows_solver_selected(int n, int precision);
size_t resolve_rows_lds_bytes(int n, int Nx, int Ny);
// Estimator index of the solver schedules: 0 phase_link EMI, 1 evd EVD/STBAS, 2 evd MLE.
inline int solve_mode_index(int module, int method) {
    return module == FRINGE_MODULE_PHASE_LINK ? 0 : (method == FRINGE_METHOD_MLE ? 2 : 1);
}

// SHP-masked neighbourhood mean (despeck.cpp:300-433); mode 0 amplitude,
// 1 interferogram, 2 interferogram with coherence.  c1, c2: rows*cols
// scratch planes; out rows [row_begin, row_end) written.
int despeck_launch(const float2* b1, const float2* b2, const uint32_t* wts, int rows, int cols, int Nx, int Ny,
                   int mode, int row_begin, int row_end, float2* c1, float2* c2, float2* out, hipStream_t s);

// Per-workgroup (sum of |z|, count of valid pixels) of one band for calamp
// (calamp.cpp:180-262); partial has calamp_partials() entries.
int calamp_partials();
int calamp_launch(const float2* z, const uint8_t* mask, long long n, double2* partial, hipStream_t s);

// 0, or the code fringe_link_run would return for these options.
int link_check(const fringe_link_tile* t);
// Workspace bytes of fringe_link_run when each Gram/solve round covers
// `chunk` pixels (a multiple of cols), and the largest such chunk that fits
// in `bytes` (0 when not one row fits).
// `area_budget` bounds the FP64 re-solve's areas above 64 dates.
size_t link_workspace_for_chunk(const fringe_link_tile* t, int chunk, size_t area_budget = kResolveAreaBudget);
int link_chunk_for_bytes(const fringe_link_tile* t, size_t bytes, size_t area_budget = kResolveAreaBudget);
// fringe_link_run with an explicit chunk (arguments already checked).
int link_run_chunked(const fringe_link_tile* t, const void* d_slc, const uint32_t* d_wts, void* d_phase,
                     float* d_tcorr, void* d_comp, void* d_work, size_t work_bytes, int chunk, void* stream,
                     size_t area_budget = kResolveAreaBudget);
void set_last_error(const std::string& what);

int coherence_launch(const LinkPlan& P, const float2* slc, const uint32_t* wts, void* gamma, int* npix,
                     hipStream_t s);
int link_launch(const LinkPlan& P, const float2* slc, const uint32_t* wts, float2* phase,
                float* tcorr, float2* comp, hipStream_t s);
size_t link_gamma_bytes_per_pixel(int n, int precision);
size_t mfma_pack_bytes(int rows, int cols, int n);
int pack_launch(const float2* slc, int rows, int cols, int n, bool f16, uint32_t* pk, hipStream_t s);
int gram_mfma_launch(const LinkPlan& P, const uint32_t* pk, const uint32_t* wts, long long first,
                     int slots, hipStream_t s);
bool link_uses_mfma(int n, int precision);
// pixel rows of one workgroup tile of the staged Gram kernel (1 when it is not used)
int gram_stage_rows(int n, int Nx, int Ny);
// static LDS bytes of the solver kernel that runs above 64 dates
size_t solve_static_lds(int precision);

}  // namespace fringe
