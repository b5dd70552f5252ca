// modules.cpp -- file-level module processors: drop-in replacements of
// evd_process (src/evd/evd.cpp:16), evd_process (src/phase_link/phase_link.cpp:17)
// and nmap_process (src/nmap/nmap.cpp:18).
//
// Inputs are opened by path (stack VRT, SHP bitmask, optional byte mask),
// processed in the reference's memsize-driven row blocks with Ny rows of
// overlap (evd.cpp:170-187, :399-443, :861-868), each block run on the GPU
// by the tile engine, and written to the same ENVI rasters, with the same
// return codes, as the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include <stdlib.h>
#include <strings.h>

#include "engine_internal.hpp"
#include "fringe_amd.h"
#include "raster.hpp"

using fringe::io::DType;
using fringe::io::EnviWriter;
using fringe::io::Raster;

namespace {

const char* nz(const char* s) { return s ? s : ""; }

int ceil_words(int Nx, int Ny) { return (int)std::ceil(((2 * Ny + 1) * (2 * Nx + 1)) / 32.0); }

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    bool alloc(size_t bytes) {
        if (bytes <= n && p) return true;
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return false;
        n = bytes;
        return true;
    }
    ~DevBuf() {
        if (p) hipFree(p);
    }
};

// The reference's row-block schedule: the first block writes [0, inysize-Ny)
// and rolls back Ny rows, middle blocks write [Ny, inysize-Ny), the last
// block writes [Ny, inysize), a single block writes everything.
struct BlockPlan {
    int yoff, inysize, first, lines, rollback;
};

int block_rows(int memsize, int blocksize, int cols, int rows, int per_pixel_units) {
    int boxes = (int)((memsize * 1.0e6) / cols) / (blocksize * per_pixel_units);
    int b = boxes * blocksize;
    if (b < blocksize) b = blocksize;
    if (b > rows) b = rows;
    return b;
}

// f(i) for i in [0, n) on up to 16 host threads; every band of a stack or
// output is a file of its own, so band reads and writes run side by side.
template <typename F> static bool parallel_for(int n, F f) {
    const int hw = std::max(1, (int)std::thread::hardware_concurrency());
    const int nt = std::max(1, std::min({n, 16, hw}));
    std::atomic<int> next{0};
    std::atomic<bool> ok{true};
    auto work = [&]() {
        for (int i = next++; i < n && ok; i = next++)
            if (!f(i)) ok = false;
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work);
    work();
    for (auto& x : th) x.join();
    return ok;
}

// Host staging memory for the PCIe copies: pinned (hipHostMalloc) when the
// runtime grants it, pageable otherwise.
template <typename T> struct HostBuf {
    T* p = nullptr;
    bool pinned = false;
    std::vector<T> v;
    void alloc(size_t count) {
        if (hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault) == hipSuccess) {
            pinned = true;
            return;
        }
        (void)hipGetLastError();
        v.resize(count);
        p = v.data();
    }
    T* data() { return p; }
    ~HostBuf() {
        if (pinned) (void)hipHostFree(p);
    }
};

bool next_block(int& yoff, int& count, int blockysize, int rows, int Ny, BlockPlan& bp) {
    if (yoff >= rows) return false;
    ++count;
    bp.yoff = yoff;
    bp.inysize = std::min(blockysize, rows - yoff);
    if (count == 1) {
        bp.first = 0;
        bp.lines = bp.inysize - Ny;
        bp.rollback = Ny;
        if (yoff + blockysize >= rows) {
            bp.lines = bp.inysize;
            bp.rollback = 0;
        }
    } else if (yoff + blockysize >= rows) {
        bp.first = Ny;
        bp.lines = bp.inysize - Ny;
        bp.rollback = 0;
    } else {
        bp.first = Ny;
        bp.lines = bp.inysize - 2 * Ny;
        bp.rollback = 0;
    }
    return true;
}

void advance(int& yoff, int blockysize, int rows, const BlockPlan& bp) {
    if (yoff + blockysize < rows) yoff += bp.lines - bp.rollback;
    else yoff = rows;
}

int link_module(const fringe_evd_options* o, int module) {
    const bool verbose = o->verbose != 0;
    const std::string method = nz(o->method);
    const int Nx = o->Nx, Ny = o->Ny;
    if (verbose) {
        std::printf("Input: %s\nWeights: %s\nOutput folder: %s\nCompressed SLC: %s/%s\n", nz(o->inputDS),
                    nz(o->wtsDS), nz(o->outputFolder), nz(o->outputCompressedSlcFolder), nz(o->compSlc));
        std::printf("Method: %s  half window %d x %d  minNeighbors %d  miniStackCount %d  bandWidth %d\n",
                    method.c_str(), Nx, Ny, o->minNeighbors, o->miniStackCount, o->bandWidth);
    }
    int meth;
    if (method == "MLE") meth = FRINGE_METHOD_MLE;
    else if (method == "EVD") meth = FRINGE_METHOD_EVD;
    else if (method == "STBAS") meth = FRINGE_METHOD_STBAS;
    else {
        std::printf("Unknown method %s\n", method.c_str());
        return FRINGE_ERR_ARGUMENT;
    }
    const int nulong = ceil_words(Nx, Ny);

    std::string err;
    auto in = fringe::io::open_raster(nz(o->inputDS), &err);
    if (!in) {
        std::printf("Cannot open stack file { %s } (%s)\nExiting with error code .... (102)\n", nz(o->inputDS), err.c_str());
        return FRINGE_ERR_OPEN_INPUT;
    }
    const int cols = in->xsize, rows = in->ysize, nbands = (int)in->bands.size();
    if (verbose) std::printf("Number of rows  = %d\nNumber of cols  = %d\nNumber of bands = %d\n", rows, cols, nbands);

    if (meth == FRINGE_METHOD_STBAS && (o->bandWidth <= 0 || o->bandWidth >= nbands - 1)) {
        std::printf("Requested STBAS bandwidth %d is not in [1, %d]\n", o->bandWidth, nbands - 2);
        return FRINGE_ERR_BANDWIDTH;
    }

    auto wts = fringe::io::open_raster(nz(o->wtsDS), &err);
    if (!wts) {
        std::printf("Could not open weights dataset {%s}\nExiting with non-zero error code ... 105\n", nz(o->wtsDS));
        return FRINGE_ERR_OPEN_WTS;
    }
    {
        int code = 0;
        if (wts->xsize != cols) code = FRINGE_ERR_WIDTH;
        if (wts->ysize != rows) code = FRINGE_ERR_LENGTH;
        if ((int)wts->bands.size() != nulong) code = FRINGE_ERR_NULONG;
        const int inNx = std::atoi(wts->metadata_item("HALFWINDOWX", "ENVI").c_str());
        const int inNy = std::atoi(wts->metadata_item("HALFWINDOWY", "ENVI").c_str());
        if (inNx != Nx || inNx == 0) code = FRINGE_ERR_HALFX;
        if (inNy != Ny || inNy == 0) code = FRINGE_ERR_HALFY;
        if (code != 0) {
            std::printf("Weights do not match the stack or the window (HALFWINDOWX=%d HALFWINDOWY=%d)\n"
                        "Exiting with error code ....(%d)\n", inNx, inNy, code);
            return code;
        }
    }

    int blockysize = block_rows(o->memsize, o->blocksize, cols, rows, nbands * 20 + 4 + nulong);
    // a block must be taller than its two halo bands, or the schedule stalls
    if (blockysize < rows && blockysize <= 2 * Ny) blockysize = std::min(rows, 2 * Ny + 1);
    if (verbose) std::printf("Block size = %d lines\n", blockysize);

    std::vector<std::string> dates(nbands);
    for (int b = 0; b < nbands; ++b) {
        dates[b] = in->bands[b].metadata_item("Date", "slc");
        if (dates[b].size() != 8) {
            std::printf("Band %d does not appear to have Date information in slc metadata domain\n", b + 1);
            return FRINGE_ERR_DATE;
        }
    }
    {
        bool isdir = false;
        if (fringe::io::path_exists(nz(o->outputFolder), &isdir)) {
            std::printf("Output folder : %s already exists\n", nz(o->outputFolder));
            return isdir ? FRINGE_ERR_OUT_DIR_EXISTS : FRINGE_ERR_OUT_IS_FILE;
        }
        if (!fringe::io::make_dir(nz(o->outputFolder))) {
            std::printf("Could not create output folder: %s\n", nz(o->outputFolder));
            return FRINGE_ERR_MKDIR;
        }
    }
    std::vector<EnviWriter> outs(nbands);
    for (int b = 0; b < nbands; ++b) {
        const std::string fname = fringe::io::join_path(nz(o->outputFolder), dates[b] + ".slc");
        if (!outs[b].create(fname, cols, rows, 1, DType::CFloat32)) {
            std::printf("Could not create output SLC: %s\n", fname.c_str());
            return FRINGE_ERR_CREATE_SLC;
        }
    }
    EnviWriter corr, comp;
    if (!corr.create(fringe::io::join_path(nz(o->outputFolder), "tcorr.bin"), cols, rows, 1, DType::Float32)) {
        std::printf("Could not create temporal correlation file\n");
        return FRINGE_ERR_CREATE_AUX;
    }
    const std::string compname = fringe::io::join_path(nz(o->outputCompressedSlcFolder), nz(o->compSlc));
    if (!comp.create(compname, cols, rows, 1, DType::CFloat32)) {
        std::printf("Could not create compressed SLC file: %s\n", compname.c_str());
        return FRINGE_ERR_CREATE_AUX;
    }

    // device buffers sized for the tallest block
    const size_t px = (size_t)blockysize * cols;
    DevBuf d_slc, d_wts, d_ph, d_tc, d_co, d_work;
    fringe_link_tile t{};
    t.rows = blockysize;
    t.cols = cols;
    t.nbands = nbands;
    t.Nx = Nx;
    t.Ny = Ny;
    t.module = module;
    t.method = meth;
    t.min_neighbors = o->minNeighbors;
    t.ref_index = o->miniStackCount - 1;
    t.bandwidth = o->bandWidth;
    t.precision = o->precision;
    t.row_begin = 0;
    t.row_end = blockysize;
    const size_t wb = fringe_link_workspace_bytes(&t);
    if (wb == 0 || !d_slc.alloc(px * nbands * 8) || !d_wts.alloc(px * nulong * 4) || !d_ph.alloc(px * nbands * 8) ||
        !d_tc.alloc(px * 4) || !d_co.alloc(px * 8) || !d_work.alloc(wb)) {
        std::printf("Cannot set up the GPU tile engine (%s)\n", fringe_last_error());
        return wb == 0 ? FRINGE_ERR_UNSUPPORTED : FRINGE_ERR_DEVICE;
    }
    hipStream_t stream;
    if (hipStreamCreate(&stream) != hipSuccess) return FRINGE_ERR_DEVICE;

    HostBuf<std::complex<float>> h_slc, h_ph, h_co;
    HostBuf<uint32_t> h_wts;
    HostBuf<float> h_tc;
    h_slc.alloc(px * nbands);
    h_ph.alloc(px * nbands);
    h_co.alloc(px);
    h_wts.alloc(px * nulong);
    h_tc.alloc(px);
    std::vector<uint32_t> wts_planes(px * nulong);
    int yoff = 0, count = 0, status = 0;
    BlockPlan bp;
    while (status == 0 && next_block(yoff, count, blockysize, rows, Ny, bp)) {
        const size_t bpx = (size_t)bp.inysize * cols;
        if (!parallel_for(nbands, [&](int b) {
                return fringe::io::read_window(*in, b + 1, 0, bp.yoff, cols, bp.inysize, h_slc.data() + (size_t)b * bpx);
            })) {
            status = FRINGE_ERR_READ_SLC;
            break;
        }
        if (!parallel_for(nulong, [&](int w) {
                return fringe::io::read_window(*wts, w + 1, 0, bp.yoff, cols, bp.inysize,
                                               wts_planes.data() + (size_t)w * bpx);
            })) {
            status = FRINGE_ERR_READ_WTS;
            break;
        }
        // band planes -> [pixel][word], in row chunks
        parallel_for(bp.inysize, [&](int y) {
            for (size_t p = (size_t)y * cols; p < (size_t)(y + 1) * cols; ++p)
                for (int w = 0; w < nulong; ++w) h_wts.data()[p * nulong + w] = wts_planes[(size_t)w * bpx + p];
            return true;
        });
        t.rows = bp.inysize;
        t.row_begin = bp.first;
        t.row_end = bp.first + bp.lines;
        if (hipMemcpyAsync(d_slc.p, h_slc.data(), bpx * nbands * 8, hipMemcpyHostToDevice, stream) != hipSuccess ||
            hipMemcpyAsync(d_wts.p, h_wts.data(), bpx * nulong * 4, hipMemcpyHostToDevice, stream) != hipSuccess) {
            status = FRINGE_ERR_DEVICE;
            break;
        }
        const int rc = fringe_link_run(&t, d_slc.p, (const uint32_t*)d_wts.p, d_ph.p, (float*)d_tc.p, d_co.p,
                                       d_work.p, d_work.n, stream);
        if (rc != 0) {
            std::printf("tile engine failed (%d): %s\n", rc, fringe_last_error());
            status = rc;
            break;
        }
        if (hipMemcpyAsync(h_ph.data(), d_ph.p, bpx * nbands * 8, hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipMemcpyAsync(h_tc.data(), d_tc.p, bpx * 4, hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipMemcpyAsync(h_co.data(), d_co.p, bpx * 8, hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess) {
            status = FRINGE_ERR_DEVICE;
            break;
        }
        const size_t off = (size_t)bp.first * cols;
        std::atomic<bool> slc_ok{true}, aux_ok{true};
        parallel_for(nbands + 2, [&](int b) {
            const int y0 = bp.yoff + bp.first;
            if (b < nbands) {
                if (!outs[b].write_rows(y0, bp.lines, h_ph.data() + (size_t)b * bpx + off)) slc_ok = false;
            } else if (b == nbands) {
                if (!corr.write_rows(y0, bp.lines, h_tc.data() + off)) aux_ok = false;
            } else if (!comp.write_rows(y0, bp.lines, h_co.data() + off)) {
                aux_ok = false;
            }
            return true;
        });
        if (!slc_ok) status = FRINGE_ERR_WRITE_SLC;
        else if (!aux_ok) status = FRINGE_ERR_WRITE_AUX;
        advance(yoff, blockysize, rows, bp);
        if (verbose) std::printf("\rprogress %5.1f%%", 100.0 * yoff / rows), std::fflush(stdout);
    }
    if (verbose) std::printf("\n");
    hipStreamDestroy(stream);
    return status;
}

// despeck_process (despeck.cpp:18-497): same checks, codes, block schedule
// and output type; the masked mean of every block runs on the GPU.
int despeck_module(const fringe_despeck_options* o) {
    const bool verbose = o->verbose != 0;
    const int Nx = o->Nx, Ny = o->Ny;
    const int nulong = ceil_words(Nx, Ny);
    std::string err;
    auto in = fringe::io::open_raster(nz(o->inputDS), &err);
    if (!in) {
        std::printf("Cannot open stack file { %s } (%s)\nExiting with error code .... (102)\n", nz(o->inputDS),
                    err.c_str());
        return 102;
    }
    const int cols = in->xsize, rows = in->ysize, nbands = (int)in->bands.size();
    if (o->band1 <= 0 || o->band1 > nbands) {
        std::printf("Master band %d outside the range of permissible bands\nExiting with error code .... (102)\n",
                    o->band1);
        return 102;
    }
    if (o->band2 > 0 && o->band2 > nbands) {
        std::printf("Slave band %d outside the range of permissible bands\nExiting with error code .... (102)\n",
                    o->band2);
        return 102;
    }
    auto wts = fringe::io::open_raster(nz(o->wtsDS), &err);
    if (!wts) {
        std::printf("Could not open weights dataset {%s}\nExiting with non-zero error code ... 105\n", nz(o->wtsDS));
        return 105;
    }
    {
        int code = 0;
        if (wts->xsize != cols) code = 106;
        if (wts->ysize != rows) code = 107;
        if ((int)wts->bands.size() != nulong) code = 108;
        const int inNx = std::atoi(wts->metadata_item("HALFWINDOWX", "ENVI").c_str());
        const int inNy = std::atoi(wts->metadata_item("HALFWINDOWY", "ENVI").c_str());
        if (inNx != Nx || inNx == 0) code = 109;
        if (inNy != Ny || inNy == 0) code = 110;
        if (code != 0) {
            std::printf("Weights do not match the input or the window\nExiting with error code ....(%d)\n", code);
            return code;
        }
    }
    // boxes per block = int(memsize 1e6 / cols) / (blocksize 4 (6 + nulong)) (despeck.cpp:153-162)
    int blockysize = block_rows(o->memsize, o->blocksize, cols, rows, 4 * (6 + nulong));
    if (blockysize < rows && blockysize <= 2 * Ny) blockysize = std::min(rows, 2 * Ny + 1);
    if (verbose) std::printf("Block size = %d lines\n", blockysize);
    const bool ifg = o->band2 > 0;
    const int mode = ifg ? (o->computeCoherence ? 2 : 1) : 0;
    EnviWriter out;
    if (!out.create(nz(o->outputDS), cols, rows, 1, ifg ? DType::CFloat32 : DType::Float32)) {
        std::printf("Could not create despecked dataset {%s}\nExiting with non-zero error code ... 104\n",
                    nz(o->outputDS));
        return 104;
    }
    const size_t px = (size_t)blockysize * cols;
    DevBuf d_b1, d_b2, d_wts, d_c1, d_c2, d_out;
    if (!d_b1.alloc(px * 8) || !d_b2.alloc(ifg ? px * 8 : 8) || !d_wts.alloc(px * nulong * 4) ||
        !d_c1.alloc(px * 8) || !d_c2.alloc(px * 8) || !d_out.alloc(px * 8))
        return FRINGE_ERR_DEVICE;
    hipStream_t stream;
    if (hipStreamCreate(&stream) != hipSuccess) return FRINGE_ERR_DEVICE;
    std::vector<std::complex<float>> h_b1(px), h_b2(ifg ? px : 1), h_out(px);
    std::vector<uint32_t> h_wts(px * nulong), band_buf(px);
    std::vector<float> h_re(px);
    int yoff = 0, count = 0, status = 0;
    BlockPlan bp;
    while (status == 0 && next_block(yoff, count, blockysize, rows, Ny, bp)) {
        const size_t bpx = (size_t)bp.inysize * cols;
        if (!fringe::io::read_window(*in, o->band1, 0, bp.yoff, cols, bp.inysize, h_b1.data()) ||
            (ifg && !fringe::io::read_window(*in, o->band2, 0, bp.yoff, cols, bp.inysize, h_b2.data()))) {
            std::printf("Error reading data at line %d\nExiting with error code .... (108)\n", bp.yoff);
            status = 108;
            break;
        }
        for (int w = 0; w < nulong && status == 0; ++w) {
            if (!fringe::io::read_window(*wts, w + 1, 0, bp.yoff, cols, bp.inysize, band_buf.data())) {
                status = 108;
                break;
            }
            for (size_t p = 0; p < bpx; ++p) h_wts[p * nulong + w] = band_buf[p];
        }
        if (status) break;
        if (hipMemcpyAsync(d_b1.p, h_b1.data(), bpx * 8, hipMemcpyHostToDevice, stream) != hipSuccess ||
            (ifg && hipMemcpyAsync(d_b2.p, h_b2.data(), bpx * 8, hipMemcpyHostToDevice, stream) != hipSuccess) ||
            hipMemcpyAsync(d_wts.p, h_wts.data(), bpx * nulong * 4, hipMemcpyHostToDevice, stream) != hipSuccess ||
            fringe::despeck_launch((const float2*)d_b1.p, (const float2*)d_b2.p, (const uint32_t*)d_wts.p,
                                   bp.inysize, cols, Nx, Ny, mode, bp.first, bp.first + bp.lines,
                                   (float2*)d_c1.p, (float2*)d_c2.p, (float2*)d_out.p, stream) != 0 ||
            hipMemcpyAsync(h_out.data(), d_out.p, bpx * 8, hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess) {
            status = FRINGE_ERR_DEVICE;
            break;
        }
        const size_t off = (size_t)bp.first * cols;
        bool ok;
        if (ifg) {
            ok = out.write_rows(bp.yoff + bp.first, bp.lines, h_out.data() + off);
        } else {   // the Float32 band takes the real part (despeck.cpp:450-455)
            for (size_t k = 0; k < (size_t)bp.lines * cols; ++k) h_re[k] = h_out[off + k].real();
            ok = out.write_rows(bp.yoff + bp.first, bp.lines, h_re.data());
        }
        if (!ok) {
            std::printf("Error writing despeck data at line %d\nExiting with error code .... (110)\n", bp.yoff);
            status = 110;
            break;
        }
        advance(yoff, blockysize, rows, bp);
    }
    hipStreamDestroy(stream);
    return status;
}

// calamp_process (calamp.cpp:10-262): same checks and codes; the per-band
// sums run on the GPU; the output is a VRT of the input whose bands carry the
// constants, formatted as the reference's std::ostream does (6 significant
// digits).
int calamp_module(const fringe_calamp_options* o) {
    const bool verbose = o->verbose != 0;
    std::string err;
    auto in = fringe::io::open_raster(nz(o->inputDS), &err);
    if (!in) {
        std::printf("Cannot open stack file { %s } (%s)\nExiting with error code .... (102)\n", nz(o->inputDS),
                    err.c_str());
        return 102;
    }
    const int cols = in->xsize, rows = in->ysize, nbands = (int)in->bands.size();
    if (verbose) std::printf("Number of rows  = %d\nNumber of cols  = %d\nNumber of bands = %d\n", rows, cols, nbands);
    std::shared_ptr<Raster> msk;
    const std::string mname = nz(o->maskDS);
    if (!mname.empty() && strcasecmp(mname.c_str(), "None") != 0) {
        msk = fringe::io::open_raster(mname, &err);
        if (!msk) {
            std::printf("Cannot open mask file { %s }\nExiting with error code .... (103)\n", mname.c_str());
            return 103;
        }
        int code = 0;
        if (msk->xsize != cols) code = 104;
        if (msk->ysize != rows) code = 105;
        if ((int)msk->bands.size() != 1) code = 106;
        if (code) {
            std::printf("Mask does not match the stack\nExiting with error code .... (%d)\n", code);
            return code;
        }
    }
    // blockysize = blocksize * (int(memsize 1e6 / cols) / (blocksize * 8)) (calamp.cpp:112-120)
    int blockysize = o->blocksize * (int((o->memsize * 1.0e6) / cols) / (o->blocksize * 8));
    if (blockysize < o->blocksize) blockysize = o->blocksize;
    if (blockysize > rows) blockysize = rows;
    if (verbose) std::printf("Block size = %d lines\n", blockysize);
    const size_t px = (size_t)blockysize * cols;
    const int np = fringe::calamp_partials();
    DevBuf d_z, d_m, d_part;
    if (!d_z.alloc(px * 8) || (msk && !d_m.alloc(px)) || !d_part.alloc((size_t)np * sizeof(double2)))
        return FRINGE_ERR_DEVICE;
    hipStream_t stream;
    if (hipStreamCreate(&stream) != hipSuccess) return FRINGE_ERR_DEVICE;
    std::vector<std::complex<float>> h_z(px);
    std::vector<uint8_t> h_m(msk ? px : 1);
    std::vector<double2> h_part(np);
    std::vector<double> total(nbands, 0.0), norms(nbands, 0.0);
    int status = 0;
    for (int yoff = 0; yoff < rows && status == 0; yoff += blockysize) {
        const int h = std::min(blockysize, rows - yoff);
        const size_t n = (size_t)h * cols;
        if (msk) {
            if (!fringe::io::read_window(*msk, 1, 0, yoff, cols, h, h_m.data())) {
                std::printf("Error reading mask at line %d\nExiting with error code .... (107)\n", yoff);
                status = 107;
                break;
            }
            if (hipMemcpyAsync(d_m.p, h_m.data(), n, hipMemcpyHostToDevice, stream) != hipSuccess) {
                status = FRINGE_ERR_DEVICE;
                break;
            }
        }
        for (int b = 0; b < nbands; ++b) {
            if (!fringe::io::read_window(*in, b + 1, 0, yoff, cols, h, h_z.data())) {
                std::printf("Error reading data from band %d at line %d\nExiting with error code .... (108)\n",
                            b + 1, yoff);
                status = 108;
                break;
            }
            if (hipMemcpyAsync(d_z.p, h_z.data(), n * 8, hipMemcpyHostToDevice, stream) != hipSuccess ||
                fringe::calamp_launch((const float2*)d_z.p, msk ? (const uint8_t*)d_m.p : nullptr, (long long)n,
                                      (double2*)d_part.p, stream) != 0 ||
                hipMemcpyAsync(h_part.data(), d_part.p, (size_t)np * sizeof(double2), hipMemcpyDeviceToHost,
                               stream) != hipSuccess ||
                hipStreamSynchronize(stream) != hipSuccess) {
                status = FRINGE_ERR_DEVICE;
                break;
            }
            double bs = 0.0, bn = 0.0;
            for (const double2& q : h_part) {
                bs += q.x;
                bn += q.y;
            }
            total[b] += bs;
            norms[b] += bn;
        }
    }
    hipStreamDestroy(stream);
    if (status) return status;
    std::vector<std::map<std::string, std::map<std::string, std::string>>> meta(nbands);
    for (int b = 0; b < nbands; ++b) {
        if (norms[b] == 0) {
            if (verbose) std::printf("No valid data found in Band %d. Set to 1.0\n", b + 1);
            total[b] = 1.0;
        } else {
            total[b] /= norms[b];
            if (verbose) std::printf("Band %d:%g\n", b + 1, total[b]);
        }
        meta[b] = in->bands[b].metadata;
        std::ostringstream strs;
        strs << total[b];
        meta[b]["slc"]["amplitudeConstant"] = strs.str();
    }
    char absbuf[4096];
    const std::string src = realpath(nz(o->inputDS), absbuf) ? std::string(absbuf) : std::string(nz(o->inputDS));
    FILE* f = std::fopen(nz(o->outputDS), "w");
    if (!f) {
        std::printf("Could not create output VRT %s\n", nz(o->outputDS));
        return 104;
    }
    const std::string text = fringe::io::vrt_band_copy(*in, src, meta);
    const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
    std::fclose(f);
    return ok ? 0 : 104;
}

}  // namespace

extern "C" {

void fringe_evd_options_init(fringe_evd_options* o) {
    std::memset(o, 0, sizeof(*o));
    o->inputDS = "";
    o->wtsDS = "";
    o->compSlc = "compslc.bin";
    o->outputCompressedSlcFolder = "";
    o->outputFolder = "";
    o->blocksize = 64;
    o->memsize = 2048;
    o->Nx = 5;
    o->Ny = 5;
    o->minNeighbors = 2;
    o->miniStackCount = 1;
    o->bandWidth = -1;
    o->method = "MLE";
    o->precision = FRINGE_PREC_MIXED;
    o->verbose = 1;
}

int fringe_evd_process(const fringe_evd_options* o) {
    return o ? link_module(o, FRINGE_MODULE_EVD) : FRINGE_ERR_ARGUMENT;
}

int fringe_phase_link_process(const fringe_evd_options* o) {
    return o ? link_module(o, FRINGE_MODULE_PHASE_LINK) : FRINGE_ERR_ARGUMENT;
}

void fringe_nmap_options_init(fringe_nmap_options* o) {
    std::memset(o, 0, sizeof(*o));
    o->inputDS = "";
    o->maskDS = "";
    o->wtsDS = "";
    o->ncountDS = "";
    o->method = "KS2";
    o->noGPU = 0;
    o->blocksize = 64;
    o->memsize = 512;
    o->Nx = 5;
    o->Ny = 5;
    o->prob = 0.05;
    o->verbose = 1;
}

int fringe_nmap_process(const fringe_nmap_options* o) {
    if (!o) return FRINGE_ERR_ARGUMENT;
    const bool verbose = o->verbose != 0;
    const std::string method = nz(o->method);
    int meth;
    if (method == "KS2") meth = 0;
    else if (method == "AD2") meth = 1;
    else {
        std::printf("Statistics method can be KS2 or AD2\nUnknown method: %s\n", method.c_str());
        return 1;   // nmap.cpp:36-42
    }
    const int Nx = o->Nx, Ny = o->Ny;
    const int nulong = ceil_words(Nx, Ny);
    std::string err;
    auto in = fringe::io::open_raster(nz(o->inputDS), &err);
    if (!in) {
        std::printf("Cannot open stack file { %s } (%s)\nExiting with error code .... (102)\n", nz(o->inputDS), err.c_str());
        return FRINGE_ERR_OPEN_INPUT;
    }
    const int cols = in->xsize, rows = in->ysize, nbands = (int)in->bands.size();
    std::shared_ptr<Raster> msk;
    const std::string mname = nz(o->maskDS);
    if (!mname.empty() && strcasecmp(mname.c_str(), "None") != 0) {
        msk = fringe::io::open_raster(mname, &err);
        if (!msk) {
            std::printf("Cannot open mask file { %s }\nExiting with error code .... (102)\n", mname.c_str());
            return FRINGE_ERR_OPEN_INPUT;
        }
        int code = 0;
        if (msk->xsize != cols) code = 104;
        if (msk->ysize != rows) code = 105;
        if (msk->bands.size() != 1) code = 106;
        if (code) {
            std::printf("Mask file does not match the stack\nExiting with error code .... (%d)\n", code);
            return code;
        }
    }
    int blockysize = block_rows(o->memsize, o->blocksize, cols, rows, 4 * (nbands + 2 + nulong));
    if (blockysize < rows && blockysize <= 2 * Ny) blockysize = std::min(rows, 2 * Ny + 1);

    std::vector<double> aconst(nbands, 0.0);
    for (int b = 0; b < nbands; ++b) {
        const std::string s = in->bands[b].metadata_item("amplitudeConstant", "slc");
        aconst[b] = s.empty() ? 0.0 : std::atof(s.c_str());
    }
    EnviWriter wcount, wwts;
    wcount.meta["HALFWINDOWX"] = std::to_string(Nx);
    wcount.meta["HALFWINDOWY"] = std::to_string(Ny);
    wwts.meta = wcount.meta;
    if (!wcount.create(nz(o->ncountDS), cols, rows, 1, DType::Int16)) {
        std::printf("Could not create count dataset {%s}\n", nz(o->ncountDS));
        return 104;
    }
    if (!wwts.create(nz(o->wtsDS), cols, rows, nulong, DType::UInt32)) {
        std::printf("Could not create weights dataset {%s}\n", nz(o->wtsDS));
        return 105;
    }
    fringe_nmap_tile t{};
    t.rows = blockysize;
    t.cols = cols;
    t.nbands = nbands;
    t.Nx = Nx;
    t.Ny = Ny;
    t.prob = o->prob;
    t.method = meth;
    const size_t px = (size_t)blockysize * cols;
    DevBuf d_slc, d_mask, d_wts, d_cnt, d_work;
    const size_t wb = fringe_nmap_workspace_bytes(&t);
    if (wb == 0 || !d_slc.alloc(px * nbands * 8) || !d_mask.alloc(px) || !d_wts.alloc(px * nulong * 4) ||
        !d_cnt.alloc(px * 4) || !d_work.alloc(wb))
        return FRINGE_ERR_DEVICE;
    hipStream_t stream;
    if (hipStreamCreate(&stream) != hipSuccess) return FRINGE_ERR_DEVICE;
    HostBuf<std::complex<float>> h_slc;
    HostBuf<uint32_t> h_wts;
    HostBuf<int32_t> h_cnt;
    h_slc.alloc(px * nbands);
    h_wts.alloc(px * nulong);
    h_cnt.alloc(px);
    std::vector<uint8_t> h_mask(px);
    std::vector<int16_t> cnt16(px);
    int yoff = 0, count = 0, status = 0;
    BlockPlan bp;
    while (status == 0 && next_block(yoff, count, blockysize, rows, Ny, bp)) {
        const size_t bpx = (size_t)bp.inysize * cols;
        if (msk && !fringe::io::read_window(*msk, 1, 0, bp.yoff, cols, bp.inysize, h_mask.data())) status = 108;
        if (status == 0 && !parallel_for(nbands, [&](int b) {
                return fringe::io::read_window(*in, b + 1, 0, bp.yoff, cols, bp.inysize, h_slc.data() + (size_t)b * bpx);
            }))
            status = 108;
        if (status) break;
        t.rows = bp.inysize;
        t.row_begin = bp.first;
        t.row_end = bp.first + bp.lines;
        if (hipMemcpyAsync(d_slc.p, h_slc.data(), bpx * nbands * 8, hipMemcpyHostToDevice, stream) != hipSuccess ||
            (msk && hipMemcpyAsync(d_mask.p, h_mask.data(), bpx, hipMemcpyHostToDevice, stream) != hipSuccess)) {
            status = FRINGE_ERR_DEVICE;
            break;
        }
        const int rc = fringe_nmap_run(&t, d_slc.p, msk ? (const uint8_t*)d_mask.p : nullptr, aconst.data(),
                                       (uint32_t*)d_wts.p, (int32_t*)d_cnt.p, d_work.p, d_work.n, stream);
        if (rc != 0) {
            status = rc;
            break;
        }
        if (hipMemcpyAsync(h_wts.data(), d_wts.p, bpx * nulong * 4, hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipMemcpyAsync(h_cnt.data(), d_cnt.p, bpx * 4, hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess) {
            status = FRINGE_ERR_DEVICE;
            break;
        }
        const size_t off = (size_t)bp.first * cols;
        for (size_t p = 0; p < (size_t)bp.lines * cols; ++p) {
            const int32_t v = h_cnt.data()[off + p];   // Int32 buffer written to an Int16 file (nmap.cpp:535-539)
            cnt16[p] = (int16_t)(v > 32767 ? 32767 : v < -32768 ? -32768 : v);
        }
        if (!wcount.write_rows(bp.yoff + bp.first, bp.lines, cnt16.data()) ||
            !wwts.write_rows(bp.yoff + bp.first, bp.lines, h_wts.data() + off * nulong))
            status = 111;
        advance(yoff, blockysize, rows, bp);
        if (verbose) std::printf("\rprogress %5.1f%%", 100.0 * yoff / rows), std::fflush(stdout);
    }
    if (verbose) std::printf("\n");
    hipStreamDestroy(stream);
    return status;
}

void fringe_despeck_options_init(fringe_despeck_options* o) {
    std::memset(o, 0, sizeof(*o));
    o->inputDS = "";
    o->wtsDS = "";
    o->outputDS = "";
    o->blocksize = 64;
    o->memsize = 512;
    o->band1 = 1;
    o->band2 = -1;
    o->Nx = 5;
    o->Ny = 5;
    o->computeCoherence = 0;
    o->verbose = 1;
}

int fringe_despeck_process(const fringe_despeck_options* o) {
    return o ? despeck_module(o) : FRINGE_ERR_ARGUMENT;
}

void fringe_calamp_options_init(fringe_calamp_options* o) {
    std::memset(o, 0, sizeof(*o));
    o->inputDS = "";
    o->maskDS = "";
    o->outputDS = "";
    o->defaultValue = 1.0;
    o->applySqrt = 0;
    o->blocksize = 128;
    o->memsize = 256;
    o->verbose = 1;
}

int fringe_calamp_process(const fringe_calamp_options* o) {
    return o ? calamp_module(o) : FRINGE_ERR_ARGUMENT;
}

}  // extern "C"
