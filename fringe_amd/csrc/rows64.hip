// rows64.hip -- launches of the FP64 register-row solver (64 < n <= 128
// dates; kernels in solve128.hpp): the plain FP64 path, and the FP64 re-solve
// of the pixels the mixed precisions did not certify, which must come out
// equal to the plain path bit for bit and therefore runs the same chain.
#include "engine_internal.hpp"
#include "link_solve.hpp"
#include "solve128.hpp"

namespace fringe {

template <int CH> static void launch_f64(const SolveArgs& a, hipStream_t s) {
    const dim3 grid((unsigned)a.slots), blk(512);
    if (a.module == FRINGE_MODULE_PHASE_LINK)
        hipLaunchKernelGGL((solve_rows_kernel<double, 4, CH, MODE_EMI, double>), grid, blk, 0, s, a);
    else if (a.method == FRINGE_METHOD_MLE)
        hipLaunchKernelGGL((solve_rows_kernel<double, 4, CH, MODE_MLE, double>), grid, blk, 0, s, a);
    else
        hipLaunchKernelGGL((solve_rows_kernel<double, 4, CH, MODE_EVD, double>), grid, blk, 0, s, a);
}

void launch_rows_solver_f64(const SolveArgs& a, hipStream_t s) {
    if (a.n <= 80) launch_f64<20>(a, s);
    else if (a.n <= 96) launch_f64<24>(a, s);
    else if (a.n <= 112) launch_f64<28>(a, s);
    else launch_f64<32>(a, s);
}

size_t resolve_rows_lds_bytes(int n, int Nx, int Ny) {
    const size_t stat = n <= 80 ? sizeof(LRows<double, 4, 20>) : n <= 96 ? sizeof(LRows<double, 4, 24>)
                        : n <= 112 ? sizeof(LRows<double, 4, 28>) : sizeof(LRows<double, 4, 32>);
    return stat + 8 * gram_f64_group_bytes(n, Nx, Ny, 16);
}

template <int CH> static void resolve_f64(const SolveArgs& a, unsigned blocks, hipStream_t s) {
    const dim3 grid(blocks), blk(512);
    const size_t lds = 8 * gram_f64_group_bytes(a.n, a.g64.Nx, a.g64.Ny, 16);   // one Gram scratch per wave
    if (a.module == FRINGE_MODULE_PHASE_LINK)
        hipLaunchKernelGGL((resolve_rows_kernel<4, CH, MODE_EMI>), grid, blk, lds, s, a);
    else if (a.method == FRINGE_METHOD_MLE)
        hipLaunchKernelGGL((resolve_rows_kernel<4, CH, MODE_MLE>), grid, blk, lds, s, a);
    else
        hipLaunchKernelGGL((resolve_rows_kernel<4, CH, MODE_EVD>), grid, blk, lds, s, a);
}

void launch_resolve_rows(const SolveArgs& a, unsigned blocks, hipStream_t s) {
    if (a.n <= 80) resolve_f64<20>(a, blocks, s);
    else if (a.n <= 96) resolve_f64<24>(a, blocks, s);
    else if (a.n <= 112) resolve_f64<28>(a, blocks, s);
    else resolve_f64<32>(a, blocks, s);
}

}  // namespace fringe
