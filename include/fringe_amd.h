/* fringe_amd.h -- C ABI of the MI355X phase-linking library (libfringe_amd.so).
 *
 * Two levels:
 *   1. Module entry points that replace the reference's per-module
 *      processors one for one (same options, defaults, files and return
 *      codes), taking plain-old-data option structs instead of the
 *      std::string-bearing evdOptions / nmapOptions.
 *   2. A tile engine on device-resident buffers (HBM pointers plus a HIP
 *      stream), used by the module entry points, the sequential estimator and
 *      the request loop.
 *
 * All pointers named d_* are device pointers; h_* are host pointers.
 * Streams are passed as void* (a hipStream_t; NULL selects the null stream).
 */
#ifndef FRINGE_AMD_H
#define FRINGE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes --------------------------------------------------------
 * 0 and 101..121 keep the meaning they have in the reference modules
 * (src/evd/evd.cpp:54-857, src/phase_link/phase_link.cpp, src/nmap/nmap.cpp).
 * 2xx are raised by this library only. */
#define FRINGE_OK 0
#define FRINGE_ERR_BANDWIDTH 101    /* STBAS bandwidth missing or too large */
#define FRINGE_ERR_OPEN_INPUT 102   /* input stack cannot be opened */
#define FRINGE_ERR_OPEN_WTS 105     /* evd: weights cannot be opened; nmap: cannot create wts */
#define FRINGE_ERR_WIDTH 106
#define FRINGE_ERR_LENGTH 107
#define FRINGE_ERR_NULONG 108
#define FRINGE_ERR_HALFX 109
#define FRINGE_ERR_HALFY 110
#define FRINGE_ERR_NMAP_WRITE 111
#define FRINGE_ERR_DATE 112
#define FRINGE_ERR_OUT_DIR_EXISTS 113
#define FRINGE_ERR_OUT_IS_FILE 114
#define FRINGE_ERR_MKDIR 115
#define FRINGE_ERR_CREATE_SLC 116
#define FRINGE_ERR_CREATE_AUX 117
#define FRINGE_ERR_READ_SLC 118
#define FRINGE_ERR_READ_WTS 119
#define FRINGE_ERR_WRITE_SLC 120
#define FRINGE_ERR_WRITE_AUX 121
#define FRINGE_ERR_ARGUMENT 201     /* invalid tile geometry or option */
#define FRINGE_ERR_UNSUPPORTED 202  /* e.g. more dates than the kernels hold */
#define FRINGE_ERR_DEVICE 203       /* HIP runtime error */
#define FRINGE_ERR_WORKSPACE 204    /* workspace too small */

/* ---- estimator selection ------------------------------------------------- */
#define FRINGE_MODULE_EVD 0         /* semantics of src/evd/evd.cpp */
#define FRINGE_MODULE_PHASE_LINK 1  /* semantics of src/phase_link/phase_link.cpp */
#define FRINGE_METHOD_MLE 0
#define FRINGE_METHOD_EVD 1
#define FRINGE_METHOD_STBAS 2
#define FRINGE_PREC_FP64 0          /* FP64 Gram, FP64 eigensolver ("plain" path) */
#define FRINGE_PREC_MIXED 1         /* fp16 hi/lo MFMA Gram (per-date power-of-two scaled),
                                       FP32 eigensolver, FP64 |Gamma| inverse */
#define FRINGE_PREC_MIXED_BF16 2    /* as MIXED with bf16 hi/lo samples (no scaling needed) */

/* ---- module level: replaces evd_process (evd.cpp:16) and
 *      evd_process (phase_link.cpp:17); fields follow evdOptions
 *      (evd.hpp:19-68, phase_link.hpp:20-69). */
typedef struct fringe_evd_options {
    const char* inputDS;                   /* stack VRT with SLCs as bands */
    const char* wtsDS;                     /* SHP bitmask written by nmap */
    const char* compSlc;                   /* name of the compressed SLC */
    const char* outputCompressedSlcFolder;
    const char* outputFolder;
    int blocksize;                         /* 64   */
    int memsize;                           /* 2048 (MB) */
    int Nx;                                /* 5 */
    int Ny;                                /* 5 */
    int minNeighbors;                      /* 2 */
    int miniStackCount;                    /* 1 */
    int bandWidth;                         /* -1 */
    const char* method;                    /* "MLE" | "EVD" | "STBAS" */
    int precision;                         /* FRINGE_PREC_* (extension; default MIXED) */
    int verbose;                           /* print options and progress */
} fringe_evd_options;

void fringe_evd_options_init(fringe_evd_options* o);
int fringe_evd_process(const fringe_evd_options* o);        /* evd.cpp:16 */
int fringe_phase_link_process(const fringe_evd_options* o); /* phase_link.cpp:17 */

/* ---- module level: replaces nmap_process (nmap.cpp:18); fields follow
 *      nmapOptions (nmap.hpp:23-65). */
typedef struct fringe_nmap_options {
    const char* inputDS;
    const char* maskDS;    /* "" or "None": no mask */
    const char* wtsDS;     /* output bitmask (ENVI BIP UInt32, nulong bands) */
    const char* ncountDS;  /* output count (ENVI Int16) */
    const char* method;    /* "KS2" | "AD2" */
    int noGPU;             /* accepted for option compatibility; this library only runs on the GPU */
    int blocksize;         /* 64  */
    int memsize;           /* 512 */
    int Nx;                /* 5 */
    int Ny;                /* 5 */
    double prob;           /* 0.05 */
    int verbose;
} fringe_nmap_options;

void fringe_nmap_options_init(fringe_nmap_options* o);
int fringe_nmap_process(const fringe_nmap_options* o);      /* nmap.cpp:18 */

/* ---- tile engine: phase linking ----------------------------------------
 * A tile is rows x cols pixels of nbands SLCs, band-planar complex64
 * d_slc[band][row][col], with the SHP bitmask d_wts[row][col][nulong]
 * (nulong = ceil((2Ny+1)(2Nx+1)/32)).  Windows are truncated at the tile
 * edges, so a tile cut from a larger image must carry Ny halo rows above and
 * below the output rows [row_begin, row_end) (the reference's row blocks do
 * the same, evd.cpp:399-443).  Every pixel of the output rows is written:
 * d_phase[band][row][col] (complex64), d_tcorr[row][col], d_comp[row][col]
 * (complex64); untouched pixels get zeros, failed pixels a negative code in
 * tcorr (evd.cpp:606-728, phase_link.cpp:547,612). */
typedef struct fringe_link_tile {
    int rows, cols, nbands;
    int Nx, Ny;
    int module;          /* FRINGE_MODULE_* */
    int method;          /* FRINGE_METHOD_* */
    int min_neighbors;   /* honoured by phase_link; evd uses 2 (evd.cpp:566) */
    int ref_index;       /* miniStackCount - 1 */
    int bandwidth;       /* STBAS */
    int precision;       /* FRINGE_PREC_* */
    int row_begin, row_end;
} fringe_link_tile;

size_t fringe_link_workspace_bytes(const fringe_link_tile* t);
/* Replaces the per-block pixel loops of evd_process (evd.cpp:512-788) and
 * phase_link's evd_process (phase_link.cpp:479-666) on a device tile.
 * The mixed precisions certify every FP32 eigenpair: a pixel whose
 * eigen-gap, smallest eigenvector component or distance from a 1e-6
 * eigenvalue threshold / positive-definiteness boundary does not guarantee
 * 5e-4 rad is recomputed end to end by the FP64 path (FP64 Gram and FP64
 * solver) in the same call, on the same stream. */
int fringe_link_run(const fringe_link_tile* t, const void* d_slc, const uint32_t* d_wts,
                    void* d_phase, float* d_tcorr, void* d_comp, void* d_work,
                    size_t work_bytes, void* stream);
/* fringe_link_run without a workspace and without any device allocation,
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  The workspace is a fixed 256 MiB device array that each GPU
 * holds for the life of the process; a tile whose workspace cannot fit it
 * (the packed records of the whole tile plus one row of coherence matrices:
 * 1024 x 1024 pixels at 30 dates fit, 2048 x 2048 do not) is refused with
 * FRINGE_ERR_WORKSPACE before any launch.  Calls on one GPU share the
 * array, so calls that overlap on one GPU (other streams, other threads)
 * race; outputs equal fringe_link_run's bit for bit. */
int fringe_link_run_noalloc(const fringe_link_tile* t, const void* d_slc, const uint32_t* d_wts,
                            void* d_phase, float* d_tcorr, void* d_comp, void* stream);
/* Pixels of the last fringe_link_run on this workspace that the FP64
 * re-solve recomputed (0 when the precision does not certify); the caller
 * synchronises the run's stream first.  No counterpart in the reference: a
 * diagnostic of the mixed-precision path.  Returns -1 on bad arguments. */
long long fringe_link_resolved_pixels(const fringe_link_tile* t, const void* d_work, size_t work_bytes);

/* ---- batched tiles: one persistent solve launch over mixed date counts ---
 * No counterpart in the reference, whose pixel loops run one stack at a time
 * (evd.cpp:512-514; sequential.py:193-240 links mini-stacks one after the
 * other).  fringe_link_batch_run links `count` independent tiles (their own
 * nbands, window, size, min_neighbors, ref_index, bandwidth; one module,
 * method and precision for the whole batch) on one stream: the coherence
 * matrices of several tiles share each round's buffer, and every round's
 * Hermitian solves of one date class (<= 32, <= 64, <= 128, > 128 dates) run
 * in ONE persistent launch that drains a queue of all tiles' pixels, whatever
 * their date count and neighbour count (tiles on the mixed precisions'
 * register-row solver, 65..128 dates, get one grid launch each).  Outputs
 * equal fringe_link_run's per tile
 * bit for bit.  Returns 202 for a batch that mixes modules, methods or
 * precisions. */
typedef struct fringe_link_batch_item {
    fringe_link_tile tile;
    const void* d_slc;
    const uint32_t* d_wts;
    void* d_phase;
    float* d_tcorr;
    void* d_comp;
} fringe_link_batch_item;
size_t fringe_link_batch_workspace_bytes(const fringe_link_batch_item* items, int count);
int fringe_link_batch_run(const fringe_link_batch_item* items, int count, void* d_work, size_t work_bytes,
                          void* stream);

/* ---- launch schedules (set by fringe_amd/autotune.py's table) ------------
 * No counterpart in the reference.  FRINGE_SCHED_GRAM_VARIANT, keyed by
 * (nbands, Nx, Ny) (nbands taken by class: <= 32, <= 64): the mixed path's
 * Gram kernel, FRINGE_GRAM_*; a variant that does not fit the window falls
 * back to the default choice.  FRINGE_SCHED_SOLVE, keyed by (nbands class,
 * estimator: 0 phase_link EMI, 1 evd EVD/STBAS, 2 evd MLE, precision): 0
 * launches one wave per unit of work, else (waves_per_cu << 8 | grab) runs a
 * persistent grid of waves_per_cu waves per compute unit that take `grab`
 * units at a time from a queue.  Unset keys keep the built-in defaults. */
#define FRINGE_SCHED_GRAM_VARIANT 0
#define FRINGE_SCHED_SOLVE 1
#define FRINGE_SCHED_GENERIC_A 2    /* (nbands class 48/64/96/128/256, precision): the solver above 32 dates
                                       keeps each row in registers, two (four up to 48 dates) threads per row
                                       in FP32, four in FP64
                                       (2, up to 128 dates; the default there, except FP32 at 49..64 dates);
                                       otherwise one row per lane up to 64 dates (0), or its work matrix in LDS
                                       (0) or global memory (1) above 64 dates */
#define FRINGE_GRAM_DEFAULT 0
#define FRINGE_GRAM_STAGED12 1      /* 12 waves, 8 x 16-pixel tiles staged in LDS */
#define FRINGE_GRAM_STAGED8 2       /* 8 waves, 4 x 16 */
#define FRINGE_GRAM_STAGED4 3       /* 4 waves, up to 4 x 8 */
#define FRINGE_GRAM_UNSTAGED 4      /* one wave per pixel, records gathered from L2 */
int fringe_schedule_set(int kind, int key0, int key1, int key2, int value);
int fringe_schedule_get(int kind, int key0, int key1, int key2);
void fringe_schedule_clear(void);

/* Coherence matrices alone (the first stage of fringe_link_run,
 * evd.cpp:530-582): for every output pixel, slot = (row - row_begin) * cols
 * + col, d_gamma[slot][i][j] = Gamma_ij (row-major, Hermitian, unit
 * diagonal) and d_npix[slot] = number of neighbours, or -1 when the pixel is
 * not linked (centre bit clear, too few neighbours).  Elements are complex64
 * when the precision uses the MFMA Gram (MIXED / MIXED_BF16 with nbands <=
 * 128) and complex128 otherwise. */
size_t fringe_coherence_workspace_bytes(const fringe_link_tile* t);
int fringe_coherence_run(const fringe_link_tile* t, const void* d_slc, const uint32_t* d_wts, void* d_gamma,
                         int* d_npix, void* d_work, size_t work_bytes, void* stream);

/* ---- tile engine: SHP map (KS2 two-sample test) -------------------------
 * d_mask (optional, uint8 [row][col]) marks usable pixels.  h_amp_const holds
 * the per-band amplitudeConstant metadata (<= 0 or NULL: missing), normalised
 * by band 1 as nmap.cpp:204-233 does.  Output rows [row_begin, row_end) of
 * d_wts[row][col][nulong] and d_count[row][col] (int32) are written, and no
 * other rows of them; every row of d_slc within Ny of the output rows is
 * read.  Each pixel pair is tested once (the reference sweep's order), in a
 * pass that writes the workspace, and a second pass assembles the masks. */
typedef struct fringe_nmap_tile {
    int rows, cols, nbands;
    int Nx, Ny;
    double prob;
    int method;          /* 0 = KS2, 1 = AD2 (nbands >= 4) */
    int row_begin, row_end;
} fringe_nmap_tile;

size_t fringe_nmap_workspace_bytes(const fringe_nmap_tile* t);
int fringe_nmap_run(const fringe_nmap_tile* t, const void* d_slc, const uint8_t* d_mask,
                    const double* h_amp_const, uint32_t* d_wts, int32_t* d_count,
                    void* d_work, size_t work_bytes, void* stream);

/* KS decision table used by the nmap kernel: h_table[k] = 1 when
 * KolmogorovProb((k/n) sqrt(n/2)) >= prob, k = 0..n (KS2sample.hpp:91-144,
 * nmap.cpp:462).  Returns the smallest distance |p_k - prob| over k, so a
 * caller can see how far every decision is from the threshold. */
double fringe_ks_table(int n, double prob, uint8_t* h_table);

/* AD2 decision used by the nmap kernel (AD2unique.hpp:125-351): a pair is
 * similar when its standardised statistic (A2_kN - 1) / sigma_N is at most
 * the returned value, the largest statistic whose interpolated p-value is
 * still >= prob.  *h_sigma (optional) receives sigma_N for n + n samples. */
double fringe_ad2_threshold(int n, double prob, double* h_sigma);

/* ---- module level: replaces despeck_process (despeck.cpp:18); fields
 *      follow despeckOptions (despeck.hpp:17-40).  band2 <= 0: despeckle the
 *      amplitude of band1 (Float32 output); otherwise the interferogram
 *      band1 x conj(band2) (CFloat32), or its coherence when
 *      computeCoherence is set. */
typedef struct fringe_despeck_options {
    const char* inputDS;   /* stack VRT */
    const char* wtsDS;     /* SHP bitmask (required, as in the reference) */
    const char* outputDS;  /* ENVI output */
    int blocksize;         /* 64 */
    int memsize;           /* 512 */
    int band1;             /* 1 */
    int band2;             /* -1 */
    int Nx;                /* 5 */
    int Ny;                /* 5 */
    int computeCoherence;  /* 0 */
    int verbose;
} fringe_despeck_options;

void fringe_despeck_options_init(fringe_despeck_options* o);
int fringe_despeck_process(const fringe_despeck_options* o);   /* despeck.cpp:18 */

/* ---- module level: replaces calamp_process (calamp.cpp:10); fields follow
 *      calampOptions (calamp.hpp:24-40).  Writes outputDS, a VRT of the
 *      input stack whose bands carry amplitudeConstant (domain "slc"), the
 *      mean |z| of the band's valid pixels (1.0 when it has none), for nmap. */
typedef struct fringe_calamp_options {
    const char* inputDS;
    const char* maskDS;    /* "" or "None": no mask */
    const char* outputDS;
    double defaultValue;   /* 1.0 (accepted; the reference does not use it) */
    int applySqrt;         /* 0   (accepted; the reference does not use it) */
    int blocksize;         /* 128 */
    int memsize;           /* 256 */
    int verbose;
} fringe_calamp_options;

void fringe_calamp_options_init(fringe_calamp_options* o);
int fringe_calamp_process(const fringe_calamp_options* o);     /* calamp.cpp:10 */

/* ---- datum adjustment of sequential mini-stacks -------------------------
 * Replaces the VRT pixel functions of adjustMiniStacks.py:157-228: wrapped
 * phases are multiplied by the mini-stack's datum phase ("mul", CFloat32),
 * unwrapped phases added ("sum", Float32), both evaluated in double and
 * rounded to the band type as GDAL does.  npx complex64 / float32 values,
 * device pointers, asynchronous on the stream. */
int fringe_adjust_wrapped(const void* d_mini, const void* d_datum, void* d_out, long long npx, void* stream);
int fringe_adjust_unwrapped(const float* d_mini, const float* d_datum, float* d_out, long long npx,
                            void* stream);

/* ---- tile server: submit / poll / complete ------------------------------
 * No counterpart in the reference, which links one file-backed stack per
 * process call (evd.cpp:16).  A server keeps the tiles of many requests
 * resident in HBM and runs them concurrently: every request names a
 * device-resident stack (band-planar complex64, kept alive by the caller
 * until the request completes), an optional device-resident SHP mask (NULL:
 * the server computes it with the KS2 nmap at prob) and the link options.
 * The server reserves the request's workspace and outputs in one arena of
 * budget_bytes allocated at creation, admits queued requests oldest first as
 * streams and arena space free up (a request passed over too often blocks
 * later ones until it fits), and launches each admitted request on its own
 * stream, so requests overlap on the device.  Outputs stay in the arena until
 * fringe_server_release.  Codes are per request: the return value of
 * fringe_server_poll / fringe_server_wait is the request's result code
 * (0, 201-204), per-pixel codes go into its tcorr as for fringe_link_run. */
typedef struct fringe_server fringe_server;

typedef struct fringe_server_request {
    fringe_link_tile tile;
    const void* d_slc;          /* [nbands][rows][cols] complex64 on the server's device */
    const uint32_t* d_wts;      /* [rows][cols][nulong], or NULL: computed by the server */
    const uint8_t* d_mask;      /* optional validity mask for the server's nmap */
    double prob;                /* KS2 significance when d_wts is NULL (nmap.hpp default 0.05) */
} fringe_server_request;

#define FRINGE_REQ_QUEUED 0
#define FRINGE_REQ_RUNNING 1
#define FRINGE_REQ_DONE 2

/* Arena bytes a request reserves while it runs (workspace + outputs). */
size_t fringe_server_request_bytes(const fringe_server_request* r);
int fringe_server_create(int device, size_t budget_bytes, int streams, fringe_server** out);
/* Queue a request (never blocks).  201 for bad arguments, 204 if the request
 * can never fit the arena. */
int fringe_server_submit(fringe_server* s, const fringe_server_request* r, long long* id);
/* *state = FRINGE_REQ_*; returns the request's code (0 while not done). */
int fringe_server_poll(fringe_server* s, long long id, int* state);
/* Block until the request is done; returns its code. */
int fringe_server_wait(fringe_server* s, long long id);
/* Device pointers of a done request's outputs (valid until release); d_wts /
 * d_count are the server-computed SHP map, NULL if the request brought one. */
int fringe_server_result(fringe_server* s, long long id, void** d_phase, float** d_tcorr,
                         void** d_comp, uint32_t** d_wts, int32_t** d_count);
/* Milliseconds from submit to admission and from submit to completion. */
int fringe_server_times(fringe_server* s, long long id, double* queued_ms, double* total_ms);
/* Free a done request's outputs. */
int fringe_server_release(fringe_server* s, long long id);
/* Run every submitted request that can still be admitted to completion, then
 * free the arena and streams.  Queued requests that cannot be admitted once
 * nothing else runs (their space is held by outputs never released) end with
 * FRINGE_ERR_WORKSPACE instead of blocking. */
int fringe_server_destroy(fringe_server* s);

/* ---- stage timing (no counterpart in the reference; for profiling) ------
 * When enabled, every engine stage (0 pack, 1 gram, 2 solve, 3 nmap_prep,
 * 4 nmap_ks) is bracketed by HIP events on the stream it runs on.
 * fringe_timing_read waits for the recorded events and returns the number
 * of stages; ms[i] and calls[i] receive the accumulated time and count. */
void fringe_timing_enable(int on);
void fringe_timing_reset(void);
int fringe_timing_read(double* ms, long long* calls, int max_stages);
const char* fringe_timing_stage_name(int stage);

/* ---- raster utilities (ENVI and GDAL VRT subsets read by the modules) ----
 * No counterpart in the reference, which reads through GDAL; the Python
 * drivers use these to size stacks and inspect outputs without GDAL.
 * band is 1-based (0 = dataset-level metadata, e.g. domain "ENVI"). */
#define FRINGE_RASTER_C64 0
#define FRINGE_RASTER_F32 1
#define FRINGE_RASTER_U32 2
#define FRINGE_RASTER_U8 3
int fringe_raster_info(const char* path, int* xsize, int* ysize, int* nbands);
int fringe_raster_metadata(const char* path, int band, const char* key, const char* domain,
                           char* out, int cap);
int fringe_raster_read(const char* path, int band, int x0, int y0, int w, int h, int type,
                       void* out);

/* ---- utilities ---------------------------------------------------------- */
int fringe_mask_words(int Nx, int Ny);
const char* fringe_version(void);
/* Last HIP error text seen by the library on this thread (or ""). */
const char* fringe_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* FRINGE_AMD_H */
