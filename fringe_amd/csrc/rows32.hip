// rows32.hip -- launches of the FP32 register-row solver (mixed precisions,
// 64 < n <= 128 dates; kernels in solve128.hpp), a translation unit of its
// own so that it compiles beside link.hip.
#include "engine_internal.hpp"
#include "link_solve.hpp"
#include "solve128.hpp"

namespace fringe {

template <int CH> static void launch_f32(const SolveArgs& a, hipStream_t s) {
    const dim3 grid((unsigned)a.slots), blk(256);
    if (a.module == FRINGE_MODULE_PHASE_LINK)
        hipLaunchKernelGGL((solve_rows_kernel<float, 2, CH, MODE_EMI, float>), grid, blk, 0, s, a);
    else if (a.method == FRINGE_METHOD_MLE)
        hipLaunchKernelGGL((solve_rows_kernel<float, 2, CH, MODE_MLE, float>), grid, blk, 0, s, a);
    else
        hipLaunchKernelGGL((solve_rows_kernel<float, 2, CH, MODE_EVD, float>), grid, blk, 0, s, a);
}

void launch_rows_solver_f32(const SolveArgs& a, hipStream_t s) {
    if (a.n <= 80) launch_f32<40>(a, s);
    else if (a.n <= 96) launch_f32<48>(a, s);
    else if (a.n <= 112) launch_f32<56>(a, s);
    else launch_f32<64>(a, s);
}

}  // namespace fringe
