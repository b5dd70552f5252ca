// schedule.cpp -- launch schedules picked by the autotuner
// (fringe_amd/autotune.py), kept per key in a process-wide table.
//
//   Gram variant, keyed by (date class, Nx, Ny): which MFMA Gram kernel and
//   tile the mixed path launches (the staged kernel on 12 waves with 8 x 16
//   pixel tiles, on 8 waves with 4 x 16, on 4 waves with at most 4 x 8, or
//   the unstaged kernel that gathers from L2).
//   Solver launch, keyed by (date class, estimator, precision): a grid with
//   one wave per unit, or a persistent grid of `waves per CU` x CUs waves
//   that take `grab` units at a time from a queue.
//   Solver layout above 32 dates, keyed by (date class 48 / 64 / 96 / 128 /
//   256, precision): the solver that keeps each row in registers, two or
//   (up to 48 dates) four threads per row in FP32, four in FP64 (2, up to
//   128 dates); otherwise,
//   up to 64 dates, the solver with one row per lane (0), and above 64 dates
//   a work matrix in LDS where it fits (0) or in global memory (1).
//
// Unset keys use the built-in defaults (the heuristics the table replaces).
#include <map>
#include <mutex>
#include <tuple>

#include "engine_internal.hpp"

namespace fringe {
namespace {

std::mutex g_mu;
std::map<std::tuple<int, int, int, int>, int> g_table;

int date_class(int n) { return n <= 32 ? 32 : (n <= 64 ? 64 : 256); }

std::tuple<int, int, int, int> key_of(int kind, int k0, int k1, int k2) {
    if (kind == FRINGE_SCHED_GRAM_VARIANT || kind == FRINGE_SCHED_SOLVE) k0 = date_class(k0);
    if (kind == FRINGE_SCHED_GENERIC_A)
        k0 = k0 <= 48 ? 48 : (k0 <= 64 ? 64 : (k0 <= 96 ? 96 : (k0 <= 128 ? 128 : 256)));
    return std::make_tuple(kind, k0, k1, k2);
}

}  // namespace

int generic_a_default(int n, int precision_key) {
    if (n <= 32 || n > 128) return 0;
    // 49..64 dates in FP32: the one-row-per-lane solver (rows of two threads
    // measured 1.35x slower there; at 33..48 rows of four threads are 1.14x
    // faster than it)
    if (precision_key != 0 && n > 48 && n <= 64) return 0;
    return kGenericARowsDefault;
}

int schedule_lookup(int kind, int k0, int k1, int k2, int fallback) {
    std::lock_guard<std::mutex> lk(g_mu);
    const auto it = g_table.find(key_of(kind, k0, k1, k2));
    return it == g_table.end() ? fallback : it->second;
}

}  // namespace fringe

using namespace fringe;

extern "C" {

int fringe_schedule_set(int kind, int key0, int key1, int key2, int value) {
    if (kind == FRINGE_SCHED_GRAM_VARIANT) {
        if (value < FRINGE_GRAM_DEFAULT || value > FRINGE_GRAM_UNSTAGED) return FRINGE_ERR_ARGUMENT;
    } else if (kind == FRINGE_SCHED_SOLVE) {
        const int waves = value >> 8, grab = value & 0xff;
        if (value != 0 && (waves < 1 || waves > 64 || grab < 1)) return FRINGE_ERR_ARGUMENT;
    } else if (kind == FRINGE_SCHED_GENERIC_A) {
        if (value < 0 || value > 2) return FRINGE_ERR_ARGUMENT;
        if (value == 2 && key0 > 128) return FRINGE_ERR_ARGUMENT;   // rows in registers: <= 128 dates
        if (value == 1 && key0 <= 64) return FRINGE_ERR_ARGUMENT;   // no work matrix up to 64 dates
    } else {
        return FRINGE_ERR_ARGUMENT;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (kind == FRINGE_SCHED_GRAM_VARIANT && value == FRINGE_GRAM_DEFAULT)
        g_table.erase(key_of(kind, key0, key1, key2));
    else
        g_table[key_of(kind, key0, key1, key2)] = value;
    return 0;
}

int fringe_schedule_get(int kind, int key0, int key1, int key2) {
    if (kind != FRINGE_SCHED_GRAM_VARIANT && kind != FRINGE_SCHED_SOLVE && kind != FRINGE_SCHED_GENERIC_A) return -1;
    if (kind == FRINGE_SCHED_GENERIC_A) return schedule_lookup(kind, key0, key1, key2, generic_a_default(key0, key1));
    return schedule_lookup(kind, key0, key1, key2, kind == FRINGE_SCHED_SOLVE ? 0 : FRINGE_GRAM_DEFAULT);
}

void fringe_schedule_clear(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_table.clear();
}

}  // extern "C"
