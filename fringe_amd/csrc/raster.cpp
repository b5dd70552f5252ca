// raster.cpp -- ENVI / VRT access for the module processors (see raster.hpp).
#include "raster.hpp"

#include <sys/stat.h>
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace fringe {
namespace io {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
static std::string lower(std::string s) {
    for (char& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}
static std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}
static std::string dirname_of(const std::string& p) {
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string(".") : (k == 0 ? std::string("/") : p.substr(0, k));
}
std::string join_path(const std::string& a, const std::string& b) {
    if (b.empty()) return a;
    if (b[0] == '/' || a.empty()) return b;
    return a.back() == '/' ? a + b : a + "/" + b;
}
bool path_exists(const std::string& p, bool* is_dir) {
    struct stat st;
    if (stat(p.c_str(), &st) != 0) return false;
    if (is_dir) *is_dir = S_ISDIR(st.st_mode);
    return true;
}
bool make_dir(const std::string& p) {
    if (p.empty()) return false;
    bool d = false;
    if (path_exists(p, &d)) return d;
    const std::string parent = dirname_of(p);
    if (parent != p && !parent.empty() && parent != "." && !path_exists(parent, nullptr)) make_dir(parent);
    return mkdir(p.c_str(), 0777) == 0;
}

int dtype_size(DType t) {
    switch (t) {
        case DType::Byte: return 1;
        case DType::Int16: case DType::UInt16: return 2;
        case DType::Int32: case DType::UInt32: case DType::Float32: return 4;
        case DType::Float64: case DType::CFloat32: return 8;
        case DType::CFloat64: return 16;
        default: return 0;
    }
}
DType dtype_from_gdal_name(const std::string& s0) {
    const std::string s = lower(s0);
    if (s == "byte") return DType::Byte;
    if (s == "int16") return DType::Int16;
    if (s == "uint16") return DType::UInt16;
    if (s == "int32") return DType::Int32;
    if (s == "uint32") return DType::UInt32;
    if (s == "float32") return DType::Float32;
    if (s == "float64") return DType::Float64;
    if (s == "cfloat32") return DType::CFloat32;
    if (s == "cfloat64") return DType::CFloat64;
    return DType::Unknown;
}
// ENVI "data type" codes
DType dtype_from_envi_code(int c) {
    switch (c) {
        case 1: return DType::Byte;
        case 2: return DType::Int16;
        case 3: return DType::Int32;
        case 4: return DType::Float32;
        case 5: return DType::Float64;
        case 6: return DType::CFloat32;
        case 9: return DType::CFloat64;
        case 12: return DType::UInt16;
        case 13: return DType::UInt32;
        default: return DType::Unknown;
    }
}
int envi_code(DType t) {
    switch (t) {
        case DType::Byte: return 1;
        case DType::Int16: return 2;
        case DType::Int32: return 3;
        case DType::Float32: return 4;
        case DType::Float64: return 5;
        case DType::CFloat32: return 6;
        case DType::CFloat64: return 9;
        case DType::UInt16: return 12;
        case DType::UInt32: return 13;
        default: return 0;
    }
}

std::string Band::metadata_item(const std::string& key, const std::string& domain) const {
    auto d = metadata.find(lower(domain));
    if (d == metadata.end()) return "";
    auto k = d->second.find(lower(key));
    return k == d->second.end() ? "" : k->second;
}
std::string Raster::metadata_item(const std::string& key, const std::string& domain) const {
    if (lower(domain) == "envi") {
        auto k = envi_meta.find(lower(key));
        return k == envi_meta.end() ? "" : k->second;
    }
    return "";
}

// ---------------------------------------------------------------------------
// minimal XML reader for VRT files: elements, attributes, text
// ---------------------------------------------------------------------------
struct XNode {
    std::string name;                          // lower case
    std::map<std::string, std::string> attr;   // lower-case keys
    std::string text;
    std::vector<XNode> kids;
    const XNode* child(const std::string& n) const {
        for (const XNode& k : kids)
            if (k.name == n) return &k;
        return nullptr;
    }
    std::string child_text(const std::string& n, const std::string& dflt = "") const {
        const XNode* c = child(n);
        return c ? trim(c->text) : dflt;
    }
    std::string at(const std::string& n, const std::string& dflt = "") const {
        auto it = attr.find(n);
        return it == attr.end() ? dflt : it->second;
    }
};

static std::string xml_unescape(const std::string& s) {
    std::string o;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '&') {
            const size_t e = s.find(';', i);
            if (e != std::string::npos) {
                const std::string ent = s.substr(i + 1, e - i - 1);
                if (ent == "lt") o += '<';
                else if (ent == "gt") o += '>';
                else if (ent == "amp") o += '&';
                else if (ent == "quot") o += '"';
                else if (ent == "apos") o += '\'';
                else o += s.substr(i, e - i + 1);
                i = e;
                continue;
            }
        }
        o += s[i];
    }
    return o;
}

static bool xml_parse(const std::string& s, size_t& i, XNode& node) {
    // at '<'
    while (i < s.size()) {
        if (s.compare(i, 4, "<!--") == 0) {
            const size_t e = s.find("-->", i);
            if (e == std::string::npos) return false;
            i = e + 3;
        } else if (s.compare(i, 2, "<?") == 0) {
            const size_t e = s.find("?>", i);
            if (e == std::string::npos) return false;
            i = e + 2;
        } else if (s[i] == '<') {
            break;
        } else {
            ++i;
        }
    }
    if (i >= s.size() || s[i] != '<') return false;
    ++i;
    size_t j = i;
    while (j < s.size() && !std::isspace((unsigned char)s[j]) && s[j] != '>' && s[j] != '/') ++j;
    node.name = lower(s.substr(i, j - i));
    i = j;
    // attributes
    while (i < s.size()) {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
        if (i >= s.size()) return false;
        if (s[i] == '/') {
            i = s.find('>', i);
            if (i == std::string::npos) return false;
            ++i;
            return true;   // self-closing
        }
        if (s[i] == '>') {
            ++i;
            break;
        }
        size_t k = i;
        while (k < s.size() && s[k] != '=' && !std::isspace((unsigned char)s[k]) && s[k] != '>') ++k;
        const std::string key = lower(s.substr(i, k - i));
        i = k;
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
        if (i < s.size() && s[i] == '=') {
            ++i;
            while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
            const char q = s[i];
            if (q == '"' || q == '\'') {
                const size_t e = s.find(q, i + 1);
                if (e == std::string::npos) return false;
                node.attr[key] = xml_unescape(s.substr(i + 1, e - i - 1));
                i = e + 1;
            }
        } else {
            node.attr[key] = "";
        }
    }
    // content
    while (i < s.size()) {
        const size_t lt = s.find('<', i);
        if (lt == std::string::npos) return false;
        node.text += xml_unescape(s.substr(i, lt - i));
        i = lt;
        if (s.compare(i, 2, "</") == 0) {
            const size_t e = s.find('>', i);
            if (e == std::string::npos) return false;
            i = e + 1;
            return true;
        }
        if (s.compare(i, 4, "<!--") == 0) {
            const size_t e = s.find("-->", i);
            if (e == std::string::npos) return false;
            i = e + 3;
            continue;
        }
        XNode kid;
        if (!xml_parse(s, i, kid)) return false;
        node.kids.push_back(std::move(kid));
    }
    return false;
}

// ---------------------------------------------------------------------------
// ENVI
// ---------------------------------------------------------------------------
static bool read_text(const std::string& p, std::string& out) {
    std::ifstream f(p, std::ios::binary);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

static bool parse_envi_header(const std::string& txt, std::map<std::string, std::string>& kv) {
    std::istringstream in(txt);
    std::string line, first;
    if (!std::getline(in, first) || lower(trim(first)).rfind("envi", 0) != 0) return false;
    std::string pending;
    while (std::getline(in, line)) {
        if (!pending.empty()) {
            pending += "\n" + line;
            if (line.find('}') == std::string::npos) continue;
            line = pending;
            pending.clear();
        } else if (line.find('{') != std::string::npos && line.find('}') == std::string::npos) {
            pending = line;
            continue;
        }
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        kv[lower(trim(line.substr(0, eq)))] = trim(line.substr(eq + 1));
    }
    return true;
}

static std::shared_ptr<Raster> open_envi(const std::string& path, std::string* err) {
    std::string hdr_txt;
    std::string hdr = path + ".hdr";
    if (!read_text(hdr, hdr_txt)) {
        const size_t dot = path.find_last_of('.');
        const size_t slash = path.find_last_of('/');
        if (dot != std::string::npos && (slash == std::string::npos || dot > slash)) {
            hdr = path.substr(0, dot) + ".hdr";
            if (!read_text(hdr, hdr_txt)) hdr_txt.clear();
        }
    }
    std::map<std::string, std::string> kv;
    if (hdr_txt.empty() || !parse_envi_header(hdr_txt, kv)) {
        if (err) *err = "no ENVI header for " + path;
        return nullptr;
    }
    auto r = std::make_shared<Raster>();
    r->path = path;
    r->envi_meta = kv;
    r->xsize = std::atoi(kv["samples"].c_str());
    r->ysize = std::atoi(kv["lines"].c_str());
    const int nb = std::max(1, std::atoi(kv["bands"].c_str()));
    const DType t = dtype_from_envi_code(std::atoi(kv["data type"].c_str()));
    const long long off = std::atoll(kv["header offset"].c_str());
    const std::string il = lower(kv.count("interleave") ? kv["interleave"] : "bsq");
    const bool msb = std::atoi(kv["byte order"].c_str()) == 1;
    const int es = dtype_size(t);
    if (r->xsize <= 0 || r->ysize <= 0 || es == 0) {
        if (err) *err = "unsupported ENVI header for " + path;
        return nullptr;
    }
    for (int b = 0; b < nb; ++b) {
        Band band;
        band.dtype = t;
        Source s;
        s.raw = true;
        s.file = path;
        s.dtype = t;
        s.msb = msb;
        if (il == "bip") {
            s.pixel_offset = (long long)es * nb;
            s.line_offset = (long long)es * nb * r->xsize;
            s.image_offset = off + (long long)es * b;
        } else if (il == "bil") {
            s.pixel_offset = es;
            s.line_offset = (long long)es * nb * r->xsize;
            s.image_offset = off + (long long)es * r->xsize * b;
        } else {
            s.pixel_offset = es;
            s.line_offset = (long long)es * r->xsize;
            s.image_offset = off + (long long)es * r->xsize * r->ysize * b;
        }
        band.sources.push_back(s);
        r->bands.push_back(std::move(band));
    }
    return r;
}

// ---------------------------------------------------------------------------
// VRT
// ---------------------------------------------------------------------------
static std::shared_ptr<Raster> open_any(const std::string& path, std::string* err, int depth);

static std::string source_path(const XNode& sf, const std::string& vrt_dir) {
    const std::string p = trim(sf.text);
    const std::string rel = sf.at("relativetovrt", "0");
    if (rel == "1" || (!p.empty() && p[0] != '/')) return join_path(vrt_dir, p);
    return p;
}

static std::shared_ptr<Raster> open_vrt(const std::string& path, const std::string& txt, std::string* err,
                                        int depth) {
    XNode root;
    size_t i = 0;
    if (!xml_parse(txt, i, root) || root.name != "vrtdataset") {
        if (err) *err = "cannot parse VRT " + path;
        return nullptr;
    }
    auto r = std::make_shared<Raster>();
    r->path = path;
    r->xsize = std::atoi(root.at("rasterxsize").c_str());
    r->ysize = std::atoi(root.at("rasterysize").c_str());
    const std::string dir = dirname_of(path);
    std::map<std::string, std::shared_ptr<Raster>> cache;
    for (const XNode& bn : root.kids) {
        if (bn.name != "vrtrasterband") continue;
        Band band;
        band.dtype = dtype_from_gdal_name(bn.at("datatype", "Byte"));
        for (const XNode& k : bn.kids) {
            if (k.name == "metadata") {
                const std::string dom = lower(k.at("domain", ""));
                for (const XNode& m : k.kids)
                    if (m.name == "mdi") band.metadata[dom][lower(m.at("key"))] = trim(m.text);
            }
        }
        if (lower(bn.at("subclass")) == "vrtrawrasterband") {
            Source s;
            s.raw = true;
            const XNode* sf = bn.child("sourcefilename");
            if (!sf) {
                if (err) *err = "raw band without SourceFilename in " + path;
                return nullptr;
            }
            s.file = source_path(*sf, dir);
            s.image_offset = std::atoll(bn.child_text("imageoffset", "0").c_str());
            s.pixel_offset = std::atoll(bn.child_text("pixeloffset", "0").c_str());
            s.line_offset = std::atoll(bn.child_text("lineoffset", "0").c_str());
            if (s.pixel_offset == 0) s.pixel_offset = dtype_size(band.dtype);
            if (s.line_offset == 0) s.line_offset = s.pixel_offset * r->xsize;
            s.msb = lower(bn.child_text("byteorder", "LSB")) == "msb";
            s.dtype = band.dtype;
            band.sources.push_back(s);
        } else {
            for (const XNode& k : bn.kids) {
                if (k.name != "simplesource" && k.name != "complexsource") continue;
                const XNode* sf = k.child("sourcefilename");
                if (!sf) continue;
                Source s;
                const std::string sp = source_path(*sf, dir);
                if (!cache.count(sp)) cache[sp] = open_any(sp, err, depth + 1);
                s.src = cache[sp];
                if (!s.src) {
                    if (err && err->empty()) *err = "cannot open source " + sp;
                    return nullptr;
                }
                s.src_band = std::atoi(k.child_text("sourceband", "1").c_str());
                if (const XNode* sr = k.child("srcrect")) {
                    s.sx = std::atof(sr->at("xoff", "0").c_str());
                    s.sy = std::atof(sr->at("yoff", "0").c_str());
                    s.sw = std::atof(sr->at("xsize", "-1").c_str());
                    s.sh = std::atof(sr->at("ysize", "-1").c_str());
                }
                if (const XNode* dr = k.child("dstrect")) {
                    s.dx = std::atof(dr->at("xoff", "0").c_str());
                    s.dy = std::atof(dr->at("yoff", "0").c_str());
                    s.dw = std::atof(dr->at("xsize", "-1").c_str());
                    s.dh = std::atof(dr->at("ysize", "-1").c_str());
                }
                if (s.sw < 0) { s.sx = 0; s.sw = s.src->xsize; }
                if (s.sh < 0) { s.sy = 0; s.sh = s.src->ysize; }
                if (s.dw < 0) { s.dx = 0; s.dw = s.sw; }
                if (s.dh < 0) { s.dy = 0; s.dh = s.sh; }
                band.sources.push_back(s);
            }
        }
        r->bands.push_back(std::move(band));
    }
    return r;
}

static std::shared_ptr<Raster> open_any(const std::string& path, std::string* err, int depth) {
    if (depth > 8) {
        if (err) *err = "VRT nesting too deep at " + path;
        return nullptr;
    }
    bool isdir = false;
    if (!path_exists(path, &isdir) || isdir) {
        if (err) *err = "cannot open " + path;
        return nullptr;
    }
    std::string head;
    {
        std::ifstream f(path, std::ios::binary);
        char buf[256] = {0};
        f.read(buf, sizeof(buf) - 1);
        head.assign(buf, (size_t)f.gcount());
    }
    if (lower(head).find("<vrtdataset") != std::string::npos) {
        std::string txt;
        if (!read_text(path, txt)) return nullptr;
        return open_vrt(path, txt, err, depth);
    }
    return open_envi(path, err);
}

std::shared_ptr<Raster> open_raster(const std::string& path, std::string* err) {
    if (err) err->clear();
    return open_any(path, err, 0);
}

// ---------------------------------------------------------------------------
// reading
// ---------------------------------------------------------------------------
template <typename T> static void convert(const unsigned char* p, DType t, bool msb, T* out);

static void swap_bytes(unsigned char* p, int n) {
    for (int i = 0; i < n / 2; ++i) std::swap(p[i], p[n - 1 - i]);
}

template <typename T> static T from_real(double v);
template <> std::complex<float> from_real(double v) { return std::complex<float>((float)v, 0.0f); }
template <> float from_real(double v) { return (float)v; }
template <> uint32_t from_real(double v) { return (uint32_t)(v < 0 ? 0 : v > 4294967295.0 ? 4294967295.0 : v); }
template <> uint8_t from_real(double v) { return (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v); }

template <typename T> static T from_complex(double re, double im);
template <> std::complex<float> from_complex(double re, double im) { return std::complex<float>((float)re, (float)im); }
template <> float from_complex(double re, double) { return (float)re; }
template <> uint32_t from_complex(double re, double) { return from_real<uint32_t>(re); }
template <> uint8_t from_complex(double re, double) { return from_real<uint8_t>(re); }

template <typename T> static T decode(const unsigned char* src, DType t, bool msb) {
    unsigned char b[16];
    const int es = dtype_size(t);
    std::memcpy(b, src, es);
    if (msb) {
        if (t == DType::CFloat32 || t == DType::CFloat64) {
            swap_bytes(b, es / 2);
            swap_bytes(b + es / 2, es / 2);
        } else {
            swap_bytes(b, es);
        }
    }
    switch (t) {
        case DType::Byte: return from_real<T>(b[0]);
        case DType::Int16: { int16_t v; std::memcpy(&v, b, 2); return from_real<T>(v); }
        case DType::UInt16: { uint16_t v; std::memcpy(&v, b, 2); return from_real<T>(v); }
        case DType::Int32: { int32_t v; std::memcpy(&v, b, 4); return from_real<T>(v); }
        case DType::UInt32: { uint32_t v; std::memcpy(&v, b, 4); return from_real<T>(v); }
        case DType::Float32: { float v; std::memcpy(&v, b, 4); return from_real<T>(v); }
        case DType::Float64: { double v; std::memcpy(&v, b, 8); return from_real<T>(v); }
        case DType::CFloat32: { float v[2]; std::memcpy(v, b, 8); return from_complex<T>(v[0], v[1]); }
        case DType::CFloat64: { double v[2]; std::memcpy(v, b, 16); return from_complex<T>(v[0], v[1]); }
        default: return from_real<T>(0.0);
    }
}

// exact passthrough of UInt32 words (no double round trip) and CFloat32
template <> uint32_t decode<uint32_t>(const unsigned char* src, DType t, bool msb) {
    if (t == DType::UInt32 && !msb) {
        uint32_t v;
        std::memcpy(&v, src, 4);
        return v;
    }
    unsigned char b[16];
    const int es = dtype_size(t);
    std::memcpy(b, src, es);
    if (msb) swap_bytes(b, es);
    switch (t) {
        case DType::UInt32: { uint32_t v; std::memcpy(&v, b, 4); return v; }
        case DType::Int32: { int32_t v; std::memcpy(&v, b, 4); return (uint32_t)v; }
        case DType::UInt16: { uint16_t v; std::memcpy(&v, b, 2); return v; }
        case DType::Int16: { int16_t v; std::memcpy(&v, b, 2); return (uint32_t)(int32_t)v; }
        case DType::Byte: return b[0];
        default: return 0u;
    }
}

// the file sample type a T is stored as without conversion
template <typename T> struct NativeType { static constexpr DType value = DType::Unknown; };
template <> struct NativeType<std::complex<float>> { static constexpr DType value = DType::CFloat32; };
template <> struct NativeType<float> { static constexpr DType value = DType::Float32; };
template <> struct NativeType<uint32_t> { static constexpr DType value = DType::UInt32; };
template <> struct NativeType<uint8_t> { static constexpr DType value = DType::Byte; };

template <typename T>
static bool read_raw(const Source& s, int x0, int y0, int w, int h, T* out, int ldo) {
    FILE* f = std::fopen(s.file.c_str(), "rb");
    if (!f) return false;
    const int es = dtype_size(s.dtype);
    bool ok = true;
    if (s.dtype == NativeType<T>::value && !s.msb && s.pixel_offset == es && es == (int)sizeof(T)) {
        // packed samples of the output type in native byte order: straight
        // into the output rows (what decode() would produce, element by element)
        for (int y = 0; y < h && ok; ++y) {
            const long long start = s.image_offset + (long long)(y0 + y) * s.line_offset + (long long)x0 * es;
            T* row = out + (size_t)y * ldo;
            if (fseeko(f, start, SEEK_SET) != 0) { ok = false; break; }
            const size_t got = std::fread(row, sizeof(T), (size_t)w, f);
            for (size_t x = got; x < (size_t)w; ++x) row[x] = T();
        }
        std::fclose(f);
        return ok;
    }
    std::vector<unsigned char> line;
    for (int y = 0; y < h && ok; ++y) {
        const long long start = s.image_offset + (long long)(y0 + y) * s.line_offset + (long long)x0 * s.pixel_offset;
        const long long span = (long long)(w - 1) * s.pixel_offset + es;
        line.resize((size_t)span);
        if (fseeko(f, start, SEEK_SET) != 0) { ok = false; break; }
        const size_t got = std::fread(line.data(), 1, (size_t)span, f);
        if ((long long)got < span) std::memset(line.data() + got, 0, (size_t)(span - got));
        for (int x = 0; x < w; ++x) out[(size_t)y * ldo + x] = decode<T>(line.data() + (size_t)x * s.pixel_offset, s.dtype, s.msb);
    }
    std::fclose(f);
    return ok;
}

template <typename T>
static bool read_window_t(const Raster& r, int b, int x0, int y0, int w, int h, T* out, int ldo) {
    if (b < 1 || b > (int)r.bands.size()) return false;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) out[(size_t)y * ldo + x] = T();
    const Band& band = r.bands[b - 1];
    for (const Source& s : band.sources) {
        if (s.raw) {
            const int xa = std::max(0, x0), ya = std::max(0, y0);
            const int xb = std::min(r.xsize, x0 + w), yb = std::min(r.ysize, y0 + h);
            if (xa >= xb || ya >= yb) continue;
            if (!read_raw<T>(s, xa, ya, xb - xa, yb - ya, out + (size_t)(ya - y0) * ldo + (xa - x0), ldo)) return false;
            continue;
        }
        // destination window covered by this source
        const int dxa = std::max(x0, (int)std::floor(s.dx)), dya = std::max(y0, (int)std::floor(s.dy));
        const int dxb = std::min(x0 + w, (int)std::ceil(s.dx + s.dw));
        const int dyb = std::min(y0 + h, (int)std::ceil(s.dy + s.dh));
        if (dxa >= dxb || dya >= dyb) continue;
        const double fx = s.sw / s.dw, fy = s.sh / s.dh;
        const bool unit = std::fabs(fx - 1.0) < 1e-12 && std::fabs(fy - 1.0) < 1e-12;
        if (unit) {
            const int sx0 = (int)std::lround(s.sx + (dxa - s.dx)), sy0 = (int)std::lround(s.sy + (dya - s.dy));
            // clip to the source raster
            const int ox = std::max(0, -sx0), oy = std::max(0, -sy0);
            const int ww = std::min(dxb - dxa, s.src->xsize - sx0) - ox;
            const int hh = std::min(dyb - dya, s.src->ysize - sy0) - oy;
            if (ww <= 0 || hh <= 0) continue;
            if (!read_window_t<T>(*s.src, s.src_band, sx0 + ox, sy0 + oy, ww, hh,
                                  out + (size_t)(dya - y0 + oy) * ldo + (dxa - x0 + ox), ldo))
                return false;
        } else {
            // nearest-neighbour resampling, pixel by pixel (rare in these stacks)
            for (int y = dya; y < dyb; ++y)
                for (int x = dxa; x < dxb; ++x) {
                    const int sx = (int)std::floor(s.sx + (x + 0.5 - s.dx) * fx);
                    const int sy = (int)std::floor(s.sy + (y + 0.5 - s.dy) * fy);
                    if (sx < 0 || sy < 0 || sx >= s.src->xsize || sy >= s.src->ysize) continue;
                    if (!read_window_t<T>(*s.src, s.src_band, sx, sy, 1, 1, out + (size_t)(y - y0) * ldo + (x - x0), 1))
                        return false;
                }
        }
    }
    return true;
}

bool read_window(const Raster& r, int b, int x0, int y0, int w, int h, std::complex<float>* out) {
    return read_window_t(r, b, x0, y0, w, h, out, w);
}
bool read_window(const Raster& r, int b, int x0, int y0, int w, int h, float* out) {
    return read_window_t(r, b, x0, y0, w, h, out, w);
}
bool read_window(const Raster& r, int b, int x0, int y0, int w, int h, uint32_t* out) {
    return read_window_t(r, b, x0, y0, w, h, out, w);
}
bool read_window(const Raster& r, int b, int x0, int y0, int w, int h, uint8_t* out) {
    return read_window_t(r, b, x0, y0, w, h, out, w);
}

// ---------------------------------------------------------------------------
// writing
// ---------------------------------------------------------------------------
static const char* gdal_name(DType t) {
    switch (t) {
        case DType::Byte: return "Byte";
        case DType::Int16: return "Int16";
        case DType::UInt16: return "UInt16";
        case DType::Int32: return "Int32";
        case DType::UInt32: return "UInt32";
        case DType::Float32: return "Float32";
        case DType::Float64: return "Float64";
        case DType::CFloat32: return "CFloat32";
        case DType::CFloat64: return "CFloat64";
        default: return "Unknown";
    }
}

bool EnviWriter::create(const std::string& p, int xs, int ys, int nb, DType t) {
    path = p;
    xsize = xs;
    ysize = ys;
    nbands = nb;
    dtype = t;
    FILE* f = std::fopen(p.c_str(), "wb");
    if (!f) return false;
    const long long total = (long long)xs * ys * nb * dtype_size(t);
    bool ok = true;
    if (total > 0) {
        ok = fseeko(f, total - 1, SEEK_SET) == 0 && std::fputc(0, f) != EOF;
    }
    std::fclose(f);
    return ok && write_header();
}

bool EnviWriter::write_rows(int y0, int h, const void* data) {
    FILE* f = std::fopen(path.c_str(), "r+b");
    if (!f) return false;
    const long long row = (long long)xsize * nbands * dtype_size(dtype);
    bool ok = fseeko(f, (long long)y0 * row, SEEK_SET) == 0 &&
              std::fwrite(data, 1, (size_t)(row * h), f) == (size_t)(row * h);
    std::fclose(f);
    return ok;
}

bool EnviWriter::write_header() const {
    FILE* f = std::fopen((path + ".hdr").c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "ENVI\ndescription = {\n%s}\n", path.c_str());
    std::fprintf(f, "samples = %d\nlines   = %d\nbands   = %d\n", xsize, ysize, nbands);
    std::fprintf(f, "header offset = 0\nfile type = ENVI Standard\ndata type = %d\n", envi_code(dtype));
    std::fprintf(f, "interleave = bip\nbyte order = 0\n");
    for (const auto& kv : meta) std::fprintf(f, "%s = %s\n", kv.first.c_str(), kv.second.c_str());
    std::fclose(f);
    return true;
}

std::string vrt_for_envi(const std::string& file, int xsize, int ysize, int nbands, DType t) {
    std::ostringstream o;
    const int es = dtype_size(t);
    o << "<VRTDataset rasterXSize=\"" << xsize << "\" rasterYSize=\"" << ysize << "\">\n";
    for (int b = 0; b < nbands; ++b) {
        o << "  <VRTRasterBand dataType=\"" << gdal_name(t) << "\" band=\"" << b + 1
          << "\" subClass=\"VRTRawRasterBand\">\n";
        o << "    <SourceFilename relativeToVRT=\"0\">" << file << "</SourceFilename>\n";
        o << "    <ImageOffset>" << (long long)es * b << "</ImageOffset>\n";
        o << "    <PixelOffset>" << (long long)es * nbands << "</PixelOffset>\n";
        o << "    <LineOffset>" << (long long)es * nbands * xsize << "</LineOffset>\n";
        o << "    <ByteOrder>LSB</ByteOrder>\n  </VRTRasterBand>\n";
    }
    o << "</VRTDataset>\n";
    return o.str();
}

static std::string xml_escape(const std::string& v) {
    std::string o;
    for (char c : v) {
        if (c == '&') o += "&amp;";
        else if (c == '<') o += "&lt;";
        else if (c == '>') o += "&gt;";
        else if (c == '"') o += "&quot;";
        else o += c;
    }
    return o;
}

std::string vrt_band_copy(const Raster& in, const std::string& src_path,
                          const std::vector<std::map<std::string, std::map<std::string, std::string>>>& meta) {
    std::ostringstream o;
    o << "<VRTDataset rasterXSize=\"" << in.xsize << "\" rasterYSize=\"" << in.ysize << "\">\n";
    for (size_t b = 0; b < in.bands.size(); ++b) {
        o << "  <VRTRasterBand dataType=\"" << gdal_name(in.bands[b].dtype) << "\" band=\"" << b + 1 << "\">\n";
        if (b < meta.size())
            for (const auto& dom : meta[b]) {
                o << "    <Metadata";
                if (!dom.first.empty()) o << " domain=\"" << xml_escape(dom.first) << "\"";
                o << ">\n";
                for (const auto& kv : dom.second)
                    o << "      <MDI key=\"" << xml_escape(kv.first) << "\">" << xml_escape(kv.second) << "</MDI>\n";
                o << "    </Metadata>\n";
            }
        o << "    <SimpleSource>\n      <SourceFilename relativeToVRT=\"0\">" << xml_escape(src_path)
          << "</SourceFilename>\n      <SourceBand>" << b + 1 << "</SourceBand>\n    </SimpleSource>\n";
        o << "  </VRTRasterBand>\n";
    }
    o << "</VRTDataset>\n";
    return o.str();
}

}  // namespace io
}  // namespace fringe
