// rows64.hip -- launches of the FP64 register-row solver (32 < n <= 128
// dates; kernels in solve128.hpp): the plain FP64 path, and the FP64 re-solve
// of the pixels the mixed precisions did not certify, which must come out
// equal to the plain path bit for bit and therefore runs the same chain.
#include "engine_internal.hpp"
#include "link_solve.hpp"
#include "solve128.hpp"

namespace fringe {

template <int CH, int RW> static void launch_f64(const SolveArgs& a, hipStream_t s) {
    const dim3 grid((unsigned)a.slots), blk(4 * RW);
    if (a.module == FRINGE_MODULE_PHASE_LINK)
        hipLaunchKernelGGL((solve_rows_kernel<double, 4, CH, RW, MODE_EMI, double>), grid, blk, 0, s, a);
    else if (a.method == FRINGE_METHOD_MLE)
        hipLaunchKernelGGL((solve_rows_kernel<double, 4, CH, RW, MODE_MLE, double>), grid, blk, 0, s, a);
    else
        hipLaunchKernelGGL((solve_rows_kernel<double, 4, CH, RW, MODE_EVD, double>), grid, blk, 0, s, a);
}

// Rows per workgroup: 48, 64, 96, 112 or 128, the fewest that hold n (a
// whole number of waves at four threads per row; 48 rows were 1.23x faster
// than 64 at 33-48 dates, 80 rows no faster than 96 at 65-80).
void launch_rows_solver_f64(const SolveArgs& a, hipStream_t s) {
    if (a.n <= 48) launch_f64<12, 48>(a, s);
    else if (a.n <= 64) launch_f64<16, 64>(a, s);
    else if (a.n <= 80) launch_f64<20, 96>(a, s);
    else if (a.n <= 96) launch_f64<24, 96>(a, s);
    else if (a.n <= 112) launch_f64<28, 112>(a, s);
    else launch_f64<32, 128>(a, s);
}

template <int CH, int RW> static size_t lds_of(int n, int Nx, int Ny) {
    using L = LRows<double, 4, CH, RW>;
    return sizeof(L) + L::NW * gram_f64_group_bytes(n, Nx, Ny, kResolveRowsCB);
}

size_t resolve_rows_lds_bytes(int n, int Nx, int Ny) {
    if (n <= 48) return lds_of<12, 48>(n, Nx, Ny);
    if (n <= 64) return lds_of<16, 64>(n, Nx, Ny);
    if (n <= 80) return lds_of<20, 96>(n, Nx, Ny);
    if (n <= 96) return lds_of<24, 96>(n, Nx, Ny);
    if (n <= 112) return lds_of<28, 112>(n, Nx, Ny);
    return lds_of<32, 128>(n, Nx, Ny);
}

template <int CH, int RW> static void resolve_f64(const SolveArgs& a, unsigned blocks, hipStream_t s) {
    constexpr int NW = LRows<double, 4, CH, RW>::NW;
    const dim3 grid(blocks), blk(4 * RW);
    const size_t lds = NW * gram_f64_group_bytes(a.n, a.g64.Nx, a.g64.Ny, kResolveRowsCB);   // one Gram scratch per wave
    if (a.module == FRINGE_MODULE_PHASE_LINK)
        hipLaunchKernelGGL((resolve_rows_kernel<4, CH, RW, MODE_EMI>), grid, blk, lds, s, a);
    else if (a.method == FRINGE_METHOD_MLE)
        hipLaunchKernelGGL((resolve_rows_kernel<4, CH, RW, MODE_MLE>), grid, blk, lds, s, a);
    else
        hipLaunchKernelGGL((resolve_rows_kernel<4, CH, RW, MODE_EVD>), grid, blk, lds, s, a);
}

void launch_resolve_rows(const SolveArgs& a, unsigned blocks, hipStream_t s) {
    if (a.n <= 48) resolve_f64<12, 48>(a, blocks, s);
    else if (a.n <= 64) resolve_f64<16, 64>(a, blocks, s);
    else if (a.n <= 80) resolve_f64<20, 96>(a, blocks, s);
    else if (a.n <= 96) resolve_f64<24, 96>(a, blocks, s);
    else if (a.n <= 112) resolve_f64<28, 112>(a, blocks, s);
    else resolve_f64<32, 128>(a, blocks, s);
}

}  // namespace fringe
