// rows32.hip -- launches of the FP32 register-row solver (mixed precisions,
// 32 < n <= 128 dates, two or four threads per row; kernels in solve128.hpp), a translation unit of its
// own so that it compiles beside link.hip.
#include "engine_internal.hpp"
#include "link_solve.hpp"
#include "solve128.hpp"

namespace fringe {

template <int CH, int RW, int Q = 2> static void launch_f32(const SolveArgs& a, hipStream_t s) {
    const dim3 grid((unsigned)a.slots), blk(Q * RW);
    if (a.module == FRINGE_MODULE_PHASE_LINK)
        hipLaunchKernelGGL((solve_rows_kernel<float, Q, CH, RW, MODE_EMI, float>), grid, blk, 0, s, a);
    else if (a.method == FRINGE_METHOD_MLE)
        hipLaunchKernelGGL((solve_rows_kernel<float, Q, CH, RW, MODE_MLE, float>), grid, blk, 0, s, a);
    else
        hipLaunchKernelGGL((solve_rows_kernel<float, Q, CH, RW, MODE_EVD, float>), grid, blk, 0, s, a);
}

// Rows per workgroup: up to 48 dates 48 rows of four threads (three waves;
// 1.1x faster than 64 rows of two), else the fewest multiple of 32 (a whole
// number of waves at two threads per row) that holds n: 96 rows instead of
// 128 took the 72-96 date solves 1.3x faster, and four threads per row above
// 48 dates were 1.4-1.6x slower.
void launch_rows_solver_f32(const SolveArgs& a, hipStream_t s) {
    if (a.n <= 48) launch_f32<12, 48, 4>(a, s);
    else if (a.n <= 64) launch_f32<32, 64>(a, s);
    else if (a.n <= 80) launch_f32<40, 96>(a, s);
    else if (a.n <= 96) launch_f32<48, 96>(a, s);
    else if (a.n <= 112) launch_f32<56, 128>(a, s);
    else launch_f32<64, 128>(a, s);
}

}  // namespace fringe
