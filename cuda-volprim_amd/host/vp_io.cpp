// vp_io.cpp — PLY scene loading, Radiance RGBE environment loading, OpenEXR
// output (C ABI in include/vp_io.h).  Plain C++17; no GPU.
//
// Conventions follow the reference loaders: src/thesis/host/utils/io/ply.cpp:20-175,
// src/thesis/host/utils/io/hdr.cpp:32-163 and src/thesis/host/utils/io/exr.cpp:35-105.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vp_io.h"

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
    g_err = msg;
    return 1;
}

bool read_file(const char* path, std::vector<unsigned char>& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    f.seekg(0, std::ios::end);
    const std::streamoff n = f.tellg();
    f.seekg(0, std::ios::beg);
    out.resize(static_cast<size_t>(n));
    if (n > 0) f.read(reinterpret_cast<char*>(out.data()), n);
    return static_cast<bool>(f) || n == 0;
}

// ---------------------------------------------------------------- PLY ----
enum class PType { I8, U8, I16, U16, I32, U32, F32, F64, Bad };

PType ptype(const std::string& t) {
    if (t == "char" || t == "int8") return PType::I8;
    if (t == "uchar" || t == "uint8") return PType::U8;
    if (t == "short" || t == "int16") return PType::I16;
    if (t == "ushort" || t == "uint16") return PType::U16;
    if (t == "int" || t == "int32") return PType::I32;
    if (t == "uint" || t == "uint32") return PType::U32;
    if (t == "float" || t == "float32") return PType::F32;
    if (t == "double" || t == "float64") return PType::F64;
    return PType::Bad;
}

size_t psize(PType t) {
    switch (t) {
        case PType::I8: case PType::U8: return 1;
        case PType::I16: case PType::U16: return 2;
        case PType::I32: case PType::U32: case PType::F32: return 4;
        case PType::F64: return 8;
        default: return 0;
    }
}

double pread(PType t, const unsigned char* p) {
    switch (t) {
        case PType::I8: { int8_t v; std::memcpy(&v, p, 1); return v; }
        case PType::U8: return p[0];
        case PType::I16: { int16_t v; std::memcpy(&v, p, 2); return v; }
        case PType::U16: { uint16_t v; std::memcpy(&v, p, 2); return v; }
        case PType::I32: { int32_t v; std::memcpy(&v, p, 4); return v; }
        case PType::U32: { uint32_t v; std::memcpy(&v, p, 4); return v; }
        case PType::F32: { float v; std::memcpy(&v, p, 4); return v; }
        case PType::F64: { double v; std::memcpy(&v, p, 8); return v; }
        default: return 0.0;
    }
}

struct PlyProp {
    std::string name;
    PType type = PType::Bad;
    bool is_list = false;
};
struct PlyElem {
    std::string name;
    uint64_t count = 0;
    std::vector<PlyProp> props;
};

const char* const kPlyProps[14] = {"x", "y", "z", "rot_0", "rot_1", "rot_2", "rot_3", "scale_0", "scale_1", "scale_2",
                                   "albedo_0", "albedo_1", "albedo_2", "sigma_t_0"};

}  // namespace

extern "C" {

const char* vp_io_last_error(void) { return g_err.c_str(); }

void vp_io_free(void* p) { std::free(p); }

int vp_io_load_ply(const char* path, float sigma_multiplier, const float albedo_override[3], vp_primitive** out,
                   uint32_t* n_out) {
    if (!path || !out || !n_out) return fail("null argument");
    std::vector<unsigned char> buf;
    if (!read_file(path, buf)) return fail(std::string("Failed to open PLY file: ") + path);
    // header
    size_t pos = 0;
    auto line = [&]() -> std::string {
        const size_t s = pos;
        while (pos < buf.size() && buf[pos] != '\n') ++pos;
        std::string l(reinterpret_cast<const char*>(buf.data() + s), pos - s);
        if (pos < buf.size()) ++pos;
        if (!l.empty() && l.back() == '\r') l.pop_back();
        return l;
    };
    if (line() != "ply") return fail(std::string("Not a PLY file: ") + path);
    std::string format;
    std::vector<PlyElem> elems;
    while (true) {
        if (pos >= buf.size()) return fail("PLY header not terminated");
        const std::string l = line();
        std::istringstream ss(l);
        std::string kw;
        ss >> kw;
        if (kw == "format") {
            ss >> format;
        } else if (kw == "element") {
            PlyElem e;
            ss >> e.name >> e.count;
            elems.push_back(e);
        } else if (kw == "property") {
            if (elems.empty()) return fail("PLY property before any element");
            std::string t;
            ss >> t;
            PlyProp pr;
            if (t == "list") {
                std::string ct, it;
                ss >> ct >> it >> pr.name;
                pr.is_list = true;
                pr.type = ptype(it);
            } else {
                pr.type = ptype(t);
                ss >> pr.name;
            }
            if (pr.type == PType::Bad) return fail("PLY property '" + pr.name + "' has an unknown type");
            elems.back().props.push_back(pr);
        } else if (kw == "end_header") {
            break;
        }
    }
    const bool ascii = format == "ascii";
    if (!ascii && format != "binary_little_endian") return fail("PLY format '" + format + "' is not supported");
    // skip elements before "vertex" (fixed-size rows only in binary files)
    size_t ei = 0;
    for (; ei < elems.size() && elems[ei].name != "vertex"; ++ei) {
        const PlyElem& e = elems[ei];
        if (ascii) {
            for (uint64_t r = 0; r < e.count; ++r) line();
        } else {
            size_t row = 0;
            for (const auto& pr : e.props) {
                if (pr.is_list) return fail("PLY element '" + e.name + "' with lists precedes 'vertex'");
                row += psize(pr.type);
            }
            pos += row * e.count;
        }
    }
    if (ei == elems.size()) return fail(std::string("PLY file '") + path + "' has no 'vertex' element");
    const PlyElem& V = elems[ei];
    int idx[14];
    for (int k = 0; k < 14; ++k) {
        idx[k] = -1;
        for (size_t q = 0; q < V.props.size(); ++q)
            if (V.props[q].name == kPlyProps[k]) idx[k] = static_cast<int>(q);
        if (idx[k] < 0) return fail(std::string("PLY file '") + path + "' missing required property '" + kPlyProps[k] + "'");
        if (V.props[idx[k]].is_list) return fail(std::string("PLY property '") + kPlyProps[k] + "' is a list");
    }
    for (const auto& pr : V.props)
        if (pr.is_list) return fail("PLY 'vertex' element with list properties is not supported");
    std::vector<size_t> off(V.props.size());
    size_t row = 0;
    for (size_t q = 0; q < V.props.size(); ++q) {
        off[q] = row;
        row += psize(V.props[q].type);
    }
    const uint64_t N = V.count;
    if (N == 0 || N > (1ull << 30)) return fail("PLY vertex count out of range");
    // Rows are decoded, converted and validated in parallel chunks, as the
    // reference does with std::execution::par (ply.cpp:97,123); the first
    // failing row (lowest index) is the one reported.
    const bool override_alb = albedo_override && albedo_override[0] >= 0.0f && albedo_override[1] >= 0.0f &&
                              albedo_override[2] >= 0.0f;
    std::vector<size_t> line_start;  // ascii: offset of each vertex row
    if (ascii) {
        line_start.resize(static_cast<size_t>(N));
        for (uint64_t r = 0; r < N; ++r) {
            if (pos >= buf.size()) return fail("PLY ascii row " + std::to_string(r) + " is short");
            line_start[r] = pos;
            line();
        }
    } else if (pos + row * N > buf.size()) {
        return fail("PLY file truncated");
    }
    auto* prims = static_cast<vp_primitive*>(std::malloc(sizeof(vp_primitive) * N));
    if (!prims) return fail("out of host memory");
    struct RowError {
        uint64_t row = ~0ull;
        std::string msg;
    };
    // decode, convert (ply.cpp:107-142) and validate (ply.cpp:146-172) rows [b, e)
    auto work = [&](uint64_t b, uint64_t e, RowError& err) {
        std::vector<double> v(V.props.size());
        for (uint64_t i = b; i < e; ++i) {
            float r[14];
            if (ascii) {
                const char* c = reinterpret_cast<const char*>(buf.data() + line_start[i]);
                const char* lim = reinterpret_cast<const char*>(buf.data() + buf.size());
                for (size_t q = 0; q < v.size(); ++q) {
                    while (c < lim && (*c == ' ' || *c == '\t' || *c == '\r')) ++c;
                    if (c >= lim || *c == '\n') {
                        err = {i, "PLY ascii row " + std::to_string(i) + " is short"};
                        return;
                    }
                    // strtod, unlike operator>>, also reads "nan" and "inf"
                    const char* tok = c;
                    while (c < lim && *c != ' ' && *c != '\t' && *c != '\r' && *c != '\n') ++c;
                    const std::string t(tok, c);
                    char* end = nullptr;
                    v[q] = std::strtod(t.c_str(), &end);
                    if (*end) {
                        err = {i, "PLY ascii row " + std::to_string(i) + ": bad number '" + t + "'"};
                        return;
                    }
                }
                for (int k = 0; k < 14; ++k) r[k] = static_cast<float>(v[idx[k]]);
            } else {
                const unsigned char* p = buf.data() + pos + i * row;
                for (int k = 0; k < 14; ++k) r[k] = static_cast<float>(pread(V.props[idx[k]].type, p + off[idx[k]]));
            }
            vp_primitive& p = prims[i];
            for (int k = 0; k < 3; ++k) p.center[k] = r[k];
            for (int k = 0; k < 4; ++k) p.quat_wxyz[k] = r[3 + k];  // normalised by vp_scene_create
            for (int k = 0; k < 3; ++k) p.scale[k] = std::exp(r[7 + k]);
            for (int k = 0; k < 3; ++k) {
                float a = override_alb ? albedo_override[k] : r[10 + k];
                if (!std::isfinite(a)) a = 0.0f;  // ply.cpp:126-137: NaN -> 0, clamp to [0, 1]
                a = a < 0.0f ? 0.0f : (a > 1.0f ? 1.0f : a);
                p.albedo[k] = a;
            }
            p.mass = r[13] * sigma_multiplier;
            const std::string id = "Primitive " + std::to_string(i) + ": ";
            const char* bad = nullptr;
            if (!(p.scale[0] > 0.0f && p.scale[1] > 0.0f && p.scale[2] > 0.0f))
                bad = "Invalid scale - all components must be > 0";
            else if (!std::isfinite(p.scale[0]) || !std::isfinite(p.scale[1]) || !std::isfinite(p.scale[2]))
                bad = "NaN/Inf in scale";
            else if (!(p.mass > 0.0f))
                bad = "Invalid optical_thickness - must be > 0";
            else if (!std::isfinite(p.mass))
                bad = "NaN/Inf in optical_thickness";
            else if (!std::isfinite(p.center[0]) || !std::isfinite(p.center[1]) || !std::isfinite(p.center[2]))
                bad = "NaN/Inf in center";
            if (bad) {
                err = {i, "Primitive " + std::to_string(i) + ": " + bad};
                return;
            }
        }
    };
    const uint64_t nthreads = std::max<uint64_t>(
        1, std::min<uint64_t>({static_cast<uint64_t>(std::max(1u, std::thread::hardware_concurrency())), 16,
                               (N + 65535) / 65536}));
    std::vector<RowError> errs(nthreads);
    {
        std::vector<std::thread> th;
        const uint64_t chunk = (N + nthreads - 1) / nthreads;
        for (uint64_t t = 0; t < nthreads; ++t) {
            const uint64_t b = t * chunk, e = std::min<uint64_t>(N, b + chunk);
            if (b < e) th.emplace_back(work, b, e, std::ref(errs[t]));
        }
        for (auto& x : th) x.join();
    }
    for (const RowError& e : errs)  // chunks in row order: the first error is the lowest row
        if (e.row != ~0ull) {
            std::free(prims);
            return fail(e.msg);
        }
    *out = prims;
    *n_out = static_cast<uint32_t>(N);
    return 0;
}

int vp_io_save_ply(const char* path, const vp_primitive* prims, uint32_t n, float sigma_multiplier) {
    if (!path || (!prims && n)) return fail("null argument");
    if (!(sigma_multiplier > 0.0f)) return fail("sigma_multiplier must be positive");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(std::string("cannot write ") + path + ": " + std::strerror(errno));
    std::fprintf(f, "ply\nformat binary_little_endian 1.0\nelement vertex %u\n", n);
    for (const char* name : kPlyProps) std::fprintf(f, "property float %s\n", name);
    std::fprintf(f, "end_header\n");
    for (uint32_t i = 0; i < n; ++i) {
        const vp_primitive& p = prims[i];
        const float row[14] = {p.center[0], p.center[1], p.center[2], p.quat_wxyz[0], p.quat_wxyz[1], p.quat_wxyz[2],
                               p.quat_wxyz[3], std::log(p.scale[0]), std::log(p.scale[1]), std::log(p.scale[2]),
                               p.albedo[0], p.albedo[1], p.albedo[2], p.mass / sigma_multiplier};
        std::fwrite(row, sizeof(float), 14, f);
    }
    const bool ok = std::fclose(f) == 0;
    return ok ? 0 : fail("write failed");
}

// ---------------------------------------------------------------- RGBE ---
int vp_io_load_hdr(const char* path, float** rgba_out, uint32_t* w_out, uint32_t* h_out) {
    if (!path || !rgba_out || !w_out || !h_out) return fail("null argument");
    std::vector<unsigned char> buf;
    if (!read_file(path, buf)) return fail(std::string("Failed to read ") + path);
    size_t pos = 0;
    auto line = [&]() -> std::string {
        const size_t s = pos;
        while (pos < buf.size() && buf[pos] != '\n') ++pos;
        std::string l(reinterpret_cast<const char*>(buf.data() + s), pos - s);
        if (pos < buf.size()) ++pos;
        return l;
    };
    const std::string magic = line();
    if (magic != "#?RADIANCE" && magic != "#?RGBE") return fail("Not a Radiance HDR file (magic '" + magic + "')");
    bool format_ok = false;
    while (pos < buf.size()) {
        const std::string l = line();
        if (l.empty()) break;
        if (l.rfind("FORMAT=", 0) == 0) {
            if (l.substr(7) != "32-bit_rle_rgbe") return fail("Unsupported HDR format '" + l.substr(7) + "'");
            format_ok = true;
        }
    }
    if (!format_ok) return fail("HDR file missing FORMAT line");
    const std::string res = line();
    int H = 0, W = 0;
    if (std::sscanf(res.c_str(), "-Y %d +X %d", &H, &W) != 2)
        return fail("Unsupported HDR axis orientation '" + res + "' (expected '-Y H +X W')");
    if (W <= 0 || H <= 0) return fail("HDR file: invalid dimensions");
    const size_t width = static_cast<size_t>(W), height = static_cast<size_t>(H);
    float* img = static_cast<float*>(std::malloc(sizeof(float) * 4 * width * height));
    if (!img) return fail("out of host memory");
    std::vector<unsigned char> ch(4 * width);
    auto to_rgba = [](unsigned char r, unsigned char g, unsigned char b, unsigned char e, float* o) {
        if (e == 0) {
            o[0] = o[1] = o[2] = 0.0f;
        } else {
            const float f = std::ldexp(1.0f, static_cast<int>(e) - (128 + 8));
            o[0] = static_cast<float>(r) * f;
            o[1] = static_cast<float>(g) * f;
            o[2] = static_cast<float>(b) * f;
        }
        o[3] = 1.0f;
    };
    for (size_t y = 0; y < height; ++y) {
        float* out = img + y * width * 4;
        if (pos + 4 > buf.size()) {
            std::free(img);
            return fail("HDR file: unexpected EOF at scanline " + std::to_string(y));
        }
        const unsigned char* p = buf.data() + pos;
        const size_t enc_w = (static_cast<size_t>(p[2]) << 8) | p[3];
        const bool new_rle = p[0] == 2 && p[1] == 2 && (p[2] & 0x80) == 0 && enc_w == width && width >= 8 &&
                             width <= 0x7fff;
        if (new_rle) {
            pos += 4;
            for (int c = 0; c < 4; ++c) {
                size_t i = 0;
                while (i < width) {
                    if (pos >= buf.size()) {
                        std::free(img);
                        return fail("HDR file: corrupt RLE");
                    }
                    const unsigned char code = buf[pos++];
                    if (code > 128) {
                        const size_t run = code & 0x7f;
                        if (i + run > width || pos >= buf.size()) {
                            std::free(img);
                            return fail("HDR file: corrupt RLE");
                        }
                        std::fill_n(&ch[c * width + i], run, buf[pos++]);
                        i += run;
                    } else {
                        if (code == 0 || i + code > width || pos + code > buf.size()) {
                            std::free(img);
                            return fail("HDR file: corrupt RLE");
                        }
                        std::memcpy(&ch[c * width + i], &buf[pos], code);
                        pos += code;
                        i += code;
                    }
                }
            }
            for (size_t x = 0; x < width; ++x)
                to_rgba(ch[x], ch[width + x], ch[2 * width + x], ch[3 * width + x], out + 4 * x);
        } else if (p[0] == 1 && p[1] == 1 && p[2] == 1) {
            std::free(img);
            return fail("HDR file uses legacy pre-2.0 RLE encoding which is not supported");
        } else {
            if (pos + width * 4 > buf.size()) {
                std::free(img);
                return fail("HDR file: unexpected EOF in raw scanline");
            }
            for (size_t x = 0; x < width; ++x) to_rgba(p[4 * x], p[4 * x + 1], p[4 * x + 2], p[4 * x + 3], out + 4 * x);
            pos += width * 4;
        }
    }
    *rgba_out = img;
    *w_out = static_cast<uint32_t>(width);
    *h_out = static_cast<uint32_t>(height);
    return 0;
}

int vp_io_save_hdr(const char* path, const float* rgba, uint32_t w, uint32_t h) {
    if (!path || !rgba || w == 0 || h == 0) return fail("bad arguments");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(std::string("cannot write ") + path);
    std::fprintf(f, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %u +X %u\n", h, w);
    std::vector<unsigned char> px(4 * static_cast<size_t>(w));
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            const float* c = rgba + 4 * (static_cast<size_t>(y) * w + x);
            const float v = std::max(c[0], std::max(c[1], c[2]));
            unsigned char* o = &px[4 * x];
            if (!(v >= 1e-32f)) {
                o[0] = o[1] = o[2] = o[3] = 0;
            } else {
                int e = 0;
                const float m = std::frexp(v, &e) * 256.0f / v;
                o[0] = static_cast<unsigned char>(std::max(0.0f, c[0] * m));
                o[1] = static_cast<unsigned char>(std::max(0.0f, c[1] * m));
                o[2] = static_cast<unsigned char>(std::max(0.0f, c[2] * m));
                o[3] = static_cast<unsigned char>(e + 128);
            }
        }
        if (w >= 8 && w <= 0x7fff) {
            // new-style scanline: four channel planes of literal runs
            const unsigned char hdr[4] = {2, 2, static_cast<unsigned char>(w >> 8), static_cast<unsigned char>(w & 255)};
            std::fwrite(hdr, 1, 4, f);
            for (int c = 0; c < 4; ++c) {
                uint32_t x = 0;
                while (x < w) {
                    const uint32_t n = std::min<uint32_t>(128u, w - x);
                    const unsigned char code = static_cast<unsigned char>(n);
                    std::fwrite(&code, 1, 1, f);
                    for (uint32_t k = 0; k < n; ++k) std::fputc(px[4 * (x + k) + c], f);
                    x += n;
                }
            }
        } else {
            std::fwrite(px.data(), 1, px.size(), f);
        }
    }
    return std::fclose(f) == 0 ? 0 : fail("write failed");
}

// ---------------------------------------------------------------- EXR ----
namespace {
void put32(std::vector<unsigned char>& b, uint32_t v) {
    for (int i = 0; i < 4; ++i) b.push_back(static_cast<unsigned char>(v >> (8 * i)));
}
void put64(std::vector<unsigned char>& b, uint64_t v) {
    for (int i = 0; i < 8; ++i) b.push_back(static_cast<unsigned char>(v >> (8 * i)));
}
void putf(std::vector<unsigned char>& b, float f) {
    uint32_t v;
    std::memcpy(&v, &f, 4);
    put32(b, v);
}
void puts_(std::vector<unsigned char>& b, const char* s) {
    while (*s) b.push_back(static_cast<unsigned char>(*s++));
    b.push_back(0);
}
void attr(std::vector<unsigned char>& b, const char* name, const char* type, uint32_t size) {
    puts_(b, name);
    puts_(b, type);
    put32(b, size);
}
uint32_t get32(const unsigned char* p) { return p[0] | (p[1] << 8) | (p[2] << 16) | (static_cast<uint32_t>(p[3]) << 24); }
}  // namespace

// Single-part scanline OpenEXR, channels B G R as FLOAT, no compression, one
// scanline per block, image rows written bottom-up (exr.cpp:41-55 flips).
int vp_io_save_exr(const char* path, const float* rgba, uint32_t w, uint32_t h) {
    if (!path || !rgba || w == 0 || h == 0) return fail("bad arguments");
    std::vector<unsigned char> b;
    b.reserve(1024 + static_cast<size_t>(w) * h * 12 + 16 * h);
    const unsigned char magic[8] = {0x76, 0x2f, 0x31, 0x01, 2, 0, 0, 0};
    b.insert(b.end(), magic, magic + 8);
    attr(b, "channels", "chlist", 3 * (2 + 16) + 1);
    for (const char* c : {"B", "G", "R"}) {
        puts_(b, c);
        put32(b, 2);  // FLOAT
        b.push_back(0);  // pLinear
        b.push_back(0);
        b.push_back(0);
        b.push_back(0);
        put32(b, 1);
        put32(b, 1);
    }
    b.push_back(0);
    attr(b, "compression", "compression", 1);
    b.push_back(0);
    attr(b, "dataWindow", "box2i", 16);
    put32(b, 0), put32(b, 0), put32(b, w - 1), put32(b, h - 1);
    attr(b, "displayWindow", "box2i", 16);
    put32(b, 0), put32(b, 0), put32(b, w - 1), put32(b, h - 1);
    attr(b, "lineOrder", "lineOrder", 1);
    b.push_back(0);
    attr(b, "pixelAspectRatio", "float", 4);
    putf(b, 1.0f);
    attr(b, "screenWindowCenter", "v2f", 8);
    putf(b, 0.0f), putf(b, 0.0f);
    attr(b, "screenWindowWidth", "float", 4);
    putf(b, 1.0f);
    b.push_back(0);  // end of header
    const size_t table = b.size();
    for (uint32_t y = 0; y < h; ++y) put64(b, 0);
    for (uint32_t y = 0; y < h; ++y) {
        const uint64_t off = b.size();
        for (int i = 0; i < 8; ++i) b[table + 8 * y + i] = static_cast<unsigned char>(off >> (8 * i));
        put32(b, y);
        put32(b, 12 * w);
        const uint32_t src = h - 1 - y;  // vertical flip
        for (int c : {2, 1, 0})
            for (uint32_t x = 0; x < w; ++x) putf(b, rgba[4 * (static_cast<size_t>(src) * w + x) + c]);
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(std::string("cannot write ") + path);
    const bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
    return (std::fclose(f) == 0 && ok) ? 0 : fail("write failed");
}

// Reads the files vp_io_save_exr writes (B G R FLOAT, uncompressed, scanline)
// and undoes the vertical flip.
int vp_io_load_exr(const char* path, float** rgba_out, uint32_t* w_out, uint32_t* h_out) {
    if (!path || !rgba_out || !w_out || !h_out) return fail("null argument");
    std::vector<unsigned char> b;
    if (!read_file(path, b)) return fail(std::string("cannot read ") + path);
    if (b.size() < 8 || b[0] != 0x76 || b[1] != 0x2f || b[2] != 0x31 || b[3] != 0x01) return fail("not an EXR file");
    size_t pos = 8;
    int xmin = 0, ymin = 0, xmax = -1, ymax = -1, nch = 0;
    std::vector<std::string> chans;
    while (pos < b.size() && b[pos] != 0) {
        std::string name(reinterpret_cast<const char*>(&b[pos]));
        pos += name.size() + 1;
        std::string type(reinterpret_cast<const char*>(&b[pos]));
        pos += type.size() + 1;
        const uint32_t size = get32(&b[pos]);
        pos += 4;
        if (name == "dataWindow") {
            xmin = static_cast<int>(get32(&b[pos]));
            ymin = static_cast<int>(get32(&b[pos + 4]));
            xmax = static_cast<int>(get32(&b[pos + 8]));
            ymax = static_cast<int>(get32(&b[pos + 12]));
        } else if (name == "compression" && b[pos] != 0) {
            return fail("compressed EXR files are not supported");
        } else if (name == "channels") {
            size_t q = pos;
            while (b[q] != 0) {
                std::string cn(reinterpret_cast<const char*>(&b[q]));
                q += cn.size() + 1;
                if (get32(&b[q]) != 2) return fail("only FLOAT channels are supported");
                q += 16;
                chans.push_back(cn);
            }
            nch = static_cast<int>(chans.size());
        }
        pos += size;
    }
    ++pos;
    const uint32_t w = static_cast<uint32_t>(xmax - xmin + 1), h = static_cast<uint32_t>(ymax - ymin + 1);
    if (xmax < xmin || ymax < ymin || nch == 0) return fail("bad EXR header");
    float* img = static_cast<float*>(std::calloc(static_cast<size_t>(w) * h * 4, sizeof(float)));
    if (!img) return fail("out of host memory");
    pos += 8 * static_cast<size_t>(h);
    for (uint32_t k = 0; k < h; ++k) {
        if (pos + 8 > b.size()) {
            std::free(img);
            return fail("EXR truncated");
        }
        const uint32_t y = get32(&b[pos]) - static_cast<uint32_t>(ymin);
        pos += 8;
        const uint32_t dst = h - 1 - y;
        for (int c = 0; c < nch; ++c) {
            const int slot = chans[c] == "R" ? 0 : chans[c] == "G" ? 1 : chans[c] == "B" ? 2 : 3;
            for (uint32_t x = 0; x < w; ++x) {
                float v;
                std::memcpy(&v, &b[pos], 4);
                pos += 4;
                img[4 * (static_cast<size_t>(dst) * w + x) + slot] = v;
            }
        }
        for (uint32_t x = 0; x < w; ++x) img[4 * (static_cast<size_t>(dst) * w + x) + 3] = 1.0f;
    }
    *rgba_out = img;
    *w_out = w;
    *h_out = h;
    return 0;
}

}  // extern "C"
