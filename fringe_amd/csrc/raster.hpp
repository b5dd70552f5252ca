// raster.hpp -- the subset of GDAL raster access the phase-linking modules
// use, without GDAL: ENVI files (binary + .hdr) and GDAL VRT files
// (VRTRawRasterBand bands, SimpleSource/ComplexSource bands with
// SrcRect/DstRect, band metadata), read as windows converted to the caller's
// type; ENVI BIP files written window by window.
#pragma once
#include <complex>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace fringe {
namespace io {

enum class DType { Byte, Int16, UInt16, Int32, UInt32, Float32, Float64, CFloat32, CFloat64, Unknown };

int dtype_size(DType t);
DType dtype_from_gdal_name(const std::string& s);   // "CFloat32", "UInt32", ...
DType dtype_from_envi_code(int code);
int envi_code(DType t);

struct Raster;

// One source feeding a band: a raw file region, or a window of another
// raster's band.
struct Source {
    // raw layout
    bool raw = false;
    std::string file;
    long long image_offset = 0, pixel_offset = 0, line_offset = 0;
    bool msb = false;
    DType dtype = DType::Unknown;
    // simple source
    std::shared_ptr<Raster> src;
    int src_band = 1;
    double sx = 0, sy = 0, sw = -1, sh = -1;   // SrcRect (sw < 0: whole)
    double dx = 0, dy = 0, dw = -1, dh = -1;   // DstRect
};

struct Band {
    DType dtype = DType::Unknown;
    std::vector<Source> sources;
    std::map<std::string, std::map<std::string, std::string>> metadata;   // domain -> key -> value
    std::string metadata_item(const std::string& key, const std::string& domain) const;
};

struct Raster {
    std::string path;
    int xsize = 0, ysize = 0;
    std::vector<Band> bands;
    std::map<std::string, std::string> envi_meta;   // ENVI header keys (lower case)
    std::string metadata_item(const std::string& key, const std::string& domain) const;
};

// Open an ENVI or VRT raster (nullptr and a message on failure).
std::shared_ptr<Raster> open_raster(const std::string& path, std::string* err);

// Read band b (1-based) window [x0, x0+w) x [y0, y0+h) into out (row-major,
// w*h elements), converting to the requested type.  Pixels not covered by
// any source are zero.  Returns false on an I/O error.
bool read_window(const Raster& r, int b, int x0, int y0, int w, int h, std::complex<float>* out);
bool read_window(const Raster& r, int b, int x0, int y0, int w, int h, float* out);
bool read_window(const Raster& r, int b, int x0, int y0, int w, int h, uint32_t* out);
bool read_window(const Raster& r, int b, int x0, int y0, int w, int h, uint8_t* out);

// ENVI writer (interleave BIP, header "<file>.hdr", little endian), created
// zero-filled at full size; rows are written in place.
struct EnviWriter {
    std::string path;
    int xsize = 0, ysize = 0, nbands = 0;
    DType dtype = DType::Unknown;
    std::map<std::string, std::string> meta;   // extra header keys
    bool create(const std::string& p, int xs, int ys, int nb, DType t);
    // rows [y0, y0+h) of all bands, data pixel-interleaved [row][col][band]
    bool write_rows(int y0, int h, const void* data);
    bool write_header() const;
};

// VRT text describing one ENVI file as a raster (what gdal_translate -of VRT
// of that file yields for our purposes).
std::string vrt_for_envi(const std::string& file, int xsize, int ysize, int nbands, DType t);

// VRT whose band b reads band b of src_path whole and carries meta[b]
// (domain -> key -> value): what GDAL's CreateCopy of the dataset followed by
// SetMetadataItem calls reads back as.
std::string vrt_band_copy(const Raster& in, const std::string& src_path,
                          const std::vector<std::map<std::string, std::map<std::string, std::string>>>& meta);

bool path_exists(const std::string& p, bool* is_dir);
bool make_dir(const std::string& p);
std::string join_path(const std::string& a, const std::string& b);

}  // namespace io
}  // namespace fringe
