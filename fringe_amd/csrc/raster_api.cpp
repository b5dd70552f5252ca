// raster_api.cpp -- C ABI over the ENVI/VRT reader for the Python drivers
// (stack assembly, output inspection) and for host-only tests.
#include <complex>
#include <cstring>
#include <string>

#include "fringe_amd.h"
#include "raster.hpp"

extern "C" {

int fringe_raster_info(const char* path, int* xsize, int* ysize, int* nbands) {
    std::string err;
    auto r = fringe::io::open_raster(path ? path : "", &err);
    if (!r) return FRINGE_ERR_OPEN_INPUT;
    if (xsize) *xsize = r->xsize;
    if (ysize) *ysize = r->ysize;
    if (nbands) *nbands = (int)r->bands.size();
    return 0;
}

int fringe_raster_metadata(const char* path, int band, const char* key, const char* domain, char* out, int cap) {
    std::string err;
    auto r = fringe::io::open_raster(path ? path : "", &err);
    if (!r) return FRINGE_ERR_OPEN_INPUT;
    std::string v;
    if (band <= 0) v = r->metadata_item(key ? key : "", domain ? domain : "");
    else if (band <= (int)r->bands.size()) v = r->bands[band - 1].metadata_item(key ? key : "", domain ? domain : "");
    else return FRINGE_ERR_ARGUMENT;
    if (out && cap > 0) {
        std::strncpy(out, v.c_str(), (size_t)cap - 1);
        out[cap - 1] = 0;
    }
    return 0;
}

int fringe_raster_read(const char* path, int band, int x0, int y0, int w, int h, int type, void* out) {
    std::string err;
    auto r = fringe::io::open_raster(path ? path : "", &err);
    if (!r) return FRINGE_ERR_OPEN_INPUT;
    if (w <= 0 || h <= 0 || !out) return FRINGE_ERR_ARGUMENT;
    bool ok = false;
    switch (type) {
        case FRINGE_RASTER_C64: ok = fringe::io::read_window(*r, band, x0, y0, w, h, (std::complex<float>*)out); break;
        case FRINGE_RASTER_F32: ok = fringe::io::read_window(*r, band, x0, y0, w, h, (float*)out); break;
        case FRINGE_RASTER_U32: ok = fringe::io::read_window(*r, band, x0, y0, w, h, (uint32_t*)out); break;
        case FRINGE_RASTER_U8: ok = fringe::io::read_window(*r, band, x0, y0, w, h, (uint8_t*)out); break;
        default: return FRINGE_ERR_ARGUMENT;
    }
    return ok ? 0 : FRINGE_ERR_READ_SLC;
}

}  // extern "C"
