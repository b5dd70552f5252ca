// noalloc.hip -- phase linking of a device tile with no device allocation:
// fringe_link_run's workspace (the list of re-solved pixels, the coherence
// matrices of a round, the packed hi/lo records) lives in one fixed device
// array, `scratch`, instead of a caller buffer.  For callers that cannot
// allocate (code under HIP graph capture, tight serving loops).
//
// The array is a file-scope __device__ object: every GPU the library runs on
// holds its own copy, at its own address, until the process ends.  Its host
// name only stands in for it, so the device address is looked up with
// hipGetSymbolAddress once per device and only that address is used.
// Calls on one GPU share the array: two calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "engine_internal.hpp"

namespace fringe {
namespace {

constexpr size_t kScratchBytes = (size_t)256 << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// Device address of this device's scratch (looked up on its first call).
int scratch_ptr(void** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return FRINGE_ERR_DEVICE;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_last_error(std::string("fringe_link_run_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return FRINGE_ERR_DEVICE;
        }
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return 0;
}

}  // namespace
}  // namespace fringe

using namespace fringe;

extern "C" int fringe_link_run_noalloc(const fringe_link_tile* t, const void* d_slc, const uint32_t* d_wts,
                                       void* d_phase, float* d_tcorr, void* d_comp, void* stream) {
    const int st = link_check(t);
    if (st != 0) return st;
    if (!d_slc || !d_wts || !d_phase || !d_tcorr || !d_comp) return FRINGE_ERR_ARGUMENT;
    // the largest round (whole rows) whose workspace fits the fixed scratch;
    // the FP64 re-solve keeps few areas (its minimum of 256 above 64 dates,
    // about 1 MiB of them up to 64)
    constexpr size_t kAreaBudget = kScratchBytes / 64;
    const int chunk = link_chunk_for_bytes(t, kScratchBytes, kAreaBudget);
    if (chunk <= 0) {
        set_last_error("fringe_link_run_noalloc: the tile needs more than the 256 MiB fixed scratch "
                       "(packed records of the whole tile plus one row of coherence matrices)");
        return FRINGE_ERR_WORKSPACE;
    }
    void* p = nullptr;
    const int rc = scratch_ptr(&p);
    if (rc != 0) return rc;
    return link_run_chunked(t, d_slc, d_wts, d_phase, d_tcorr, d_comp, p, kScratchBytes, chunk, stream, kAreaBudget);
}
