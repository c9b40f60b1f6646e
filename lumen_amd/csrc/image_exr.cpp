// image_exr.cpp — OpenEXR single-part scanline files for the output image and
// the ground truth of the RMSE metric.
//
// Writing replaces ImageUtils::save_exr (src/Framework/ImageUtils.cpp:23-89):
// channels B, G, R (in that order, the alphabetical order the format asks
// for), the RGB of an RGBA32F buffer, rows top to bottom (increasing y), no
// compression (tinyexr's header default), stored as HALF as the reference does
// (ImageUtils.cpp:74) or as FLOAT.  Reading replaces ImageUtils::load_exr
// (ImageUtils.cpp:9-21), which loads the ground truth of the RMSE pass
// (src/RayTracer/RayTracer.cpp:199-242): any subset of the channels R, G, B, A
// in HALF, FLOAT or UINT, compressed with NONE, ZIPS (one line per chunk) or
// ZIP (16 lines per chunk), returned as RGBA32F (missing colour channels 0,
// missing alpha 1).
//
// File layout (OpenEXR file format, "Technical Introduction" and "OpenEXR
// File Layout"): magic 20000630, version word 2 (single-part scanline),
// attributes (name\0 type\0 int32 size, value) ended by a zero byte, a table of
// uint64 chunk offsets, then per chunk int32 y, int32 byte count and the
// lines of the chunk, each line channel by channel.  ZIP data is zlib-deflated
// after two filters: bytes are split into the even-indexed half followed by the
// odd-indexed half, then each byte is replaced by the difference to its
// predecessor plus 128 (mod 256).
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "lt.h"

namespace {

constexpr uint32_t kMagic = 20000630u;
enum : int32_t { kUint = 0, kHalf = 1, kFloat = 2 };
enum : uint8_t { kNone = 0, kZips = 2, kZip = 3 };

lt_status fail(char* err, size_t len, lt_status st, const char* fmt, ...) {
    if (err && len) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(err, len, fmt, ap);
        va_end(ap);
    }
    return st;
}

// float -> binary16, round to nearest even (IEEE 754-2008 convertFormat);
// overflow to infinity, NaN kept quiet with its top payload bits.
uint16_t f32_to_f16(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u)  // inf or NaN
        return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u | ((ax >> 13) & 0x3ffu) : 0u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // rounds to >= 65520 -> inf
    if (ax < 0x38800000u) {                                     // below 2^-14: subnormal or zero
        if (ax < 0x33000000u) return (uint16_t)sign;             // below 2^-25: rounds to 0
        const uint32_t e = ax >> 23, m = (ax & 0x7fffffu) | 0x800000u;
        const uint32_t shift = 126u - e;  // 14..24 for e in 102..112
        uint32_t h = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((ax >> 13) - (112u << 10));  // rebias 127 -> 15
    const uint32_t rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;  // may carry into the exponent (exact)
    return (uint16_t)(sign | h);
}

float f16_to_f32(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu, x;
    if (e == 0) {
        if (m == 0) {
            x = sign;
        } else {  // subnormal: normalise
            e = 113;
            while (!(m & 0x400u)) {
                m <<= 1;
                e--;
            }
            x = sign | (e << 23) | ((m & 0x3ffu) << 13);
        }
    } else if (e == 31) {
        x = sign | 0x7f800000u | (m << 13);
    } else {
        x = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

void put_u32(std::vector<uint8_t>& b, uint32_t v) {
    for (int i = 0; i < 4; i++) b.push_back((uint8_t)(v >> (8 * i)));
}
void put_u64(std::vector<uint8_t>& b, uint64_t v) {
    for (int i = 0; i < 8; i++) b.push_back((uint8_t)(v >> (8 * i)));
}
void put_f32(std::vector<uint8_t>& b, float f) {
    uint32_t v;
    memcpy(&v, &f, 4);
    put_u32(b, v);
}
void put_str(std::vector<uint8_t>& b, const char* s) {
    b.insert(b.end(), s, s + strlen(s) + 1);
}
void put_attr(std::vector<uint8_t>& b, const char* name, const char* type, const std::vector<uint8_t>& v) {
    put_str(b, name);
    put_str(b, type);
    put_u32(b, (uint32_t)v.size());
    b.insert(b.end(), v.begin(), v.end());
}

struct Reader {
    const uint8_t* p;
    size_t n, at = 0;
    bool ok = true;
    bool need(size_t k) {
        if (n - at < k || at > n) ok = false;
        return ok;
    }
    uint32_t u32() {
        if (!need(4)) return 0;
        uint32_t v = (uint32_t)p[at] | (uint32_t)p[at + 1] << 8 | (uint32_t)p[at + 2] << 16 | (uint32_t)p[at + 3] << 24;
        at += 4;
        return v;
    }
    uint64_t u64() {
        uint64_t lo = u32();
        return lo | (uint64_t)u32() << 32;
    }
    std::string cstr() {
        size_t e = at;
        while (e < n && p[e] != 0) e++;
        if (e >= n) {
            ok = false;
            return {};
        }
        std::string s((const char*)p + at, e - at);
        at = e + 1;
        return s;
    }
};

struct Channel {
    std::string name;
    int32_t type;
    int slot;  // 0..3 = R, G, B, A; -1 = ignored
};

uint32_t type_bytes(int32_t t) { return t == kHalf ? 2u : 4u; }

// undo the ZIP filters of one decompressed chunk (predictor, then interleave)
void unfilter(std::vector<uint8_t>& buf, std::vector<uint8_t>& tmp) {
    const size_t n = buf.size();
    for (size_t i = 1; i < n; i++) buf[i] = (uint8_t)(buf[i - 1] + buf[i] - 128);
    tmp.resize(n);
    const size_t h = (n + 1) / 2;
    for (size_t i = 0; i < n; i++) tmp[i] = (i & 1) ? buf[h + i / 2] : buf[i / 2];
    buf.swap(tmp);
}

}  // namespace

static lt_status exr_save_impl(const char* path, const float* rgba, uint32_t width, uint32_t height,
                                 int32_t pixel_type, char* err, size_t err_len) {
    if (!path || !rgba || width == 0 || height == 0 || width > (1u << 30) / height)
        return fail(err, err_len, LT_ERR_INVALID_ARG, "lt_exr_save: null pointer or bad size");
    if (pixel_type != LT_EXR_HALF && pixel_type != LT_EXR_FLOAT)
        return fail(err, err_len, LT_ERR_INVALID_ARG, "lt_exr_save: pixel type %d (HALF = 1, FLOAT = 2)", pixel_type);
    const uint32_t bpc = type_bytes(pixel_type);
    std::vector<uint8_t> f;
    put_u32(f, kMagic);
    put_u32(f, 2u);
    std::vector<uint8_t> v;
    for (const char* c : {"B", "G", "R"}) {
        put_str(v, c);
        put_u32(v, (uint32_t)pixel_type);
        put_u32(v, 0u);  // pLinear 0, three reserved bytes
        put_u32(v, 1u);  // x sampling
        put_u32(v, 1u);  // y sampling
    }
    v.push_back(0);
    put_attr(f, "channels", "chlist", v);
    put_attr(f, "compression", "compression", {kNone});
    v.clear();
    for (uint32_t x : {0u, 0u, width - 1u, height - 1u}) put_u32(v, x);
    put_attr(f, "dataWindow", "box2i", v);
    put_attr(f, "displayWindow", "box2i", v);
    put_attr(f, "lineOrder", "lineOrder", {0});  // INCREASING_Y
    v.clear();
    put_f32(v, 1.f);
    put_attr(f, "pixelAspectRatio", "float", v);
    v.clear();
    put_f32(v, 0.f);
    put_f32(v, 0.f);
    put_attr(f, "screenWindowCenter", "v2f", v);
    v.clear();
    put_f32(v, 1.f);
    put_attr(f, "screenWindowWidth", "float", v);
    f.push_back(0);  // end of header
    const uint64_t line_bytes = 3ull * width * bpc;
    const uint64_t table = f.size(), first = table + 8ull * height;
    for (uint32_t y = 0; y < height; y++) put_u64(f, first + (uint64_t)y * (8 + line_bytes));
    f.reserve(first + (uint64_t)height * (8 + line_bytes));
    for (uint32_t y = 0; y < height; y++) {
        put_u32(f, y);
        put_u32(f, (uint32_t)line_bytes);
        for (int c : {2, 1, 0})
            for (uint32_t x = 0; x < width; x++) {
                const float val = rgba[4 * ((size_t)y * width + x) + c];
                if (pixel_type == LT_EXR_HALF) {
                    const uint16_t h = f32_to_f16(val);
                    f.push_back((uint8_t)h);
                    f.push_back((uint8_t)(h >> 8));
                } else {
                    put_f32(f, val);
                }
            }
    }
    FILE* fp = fopen(path, "wb");
    if (!fp) return fail(err, err_len, LT_ERR_INVALID_ARG, "lt_exr_save: cannot open %s", path);
    const size_t wrote = fwrite(f.data(), 1, f.size(), fp);
    const int closed = fclose(fp);
    if (wrote != f.size() || closed != 0)
        return fail(err, err_len, LT_ERR_INVALID_ARG, "lt_exr_save: short write to %s", path);
    return LT_OK;
}

extern "C" lt_status lt_exr_save(const char* path, const float* rgba, uint32_t width, uint32_t height,
                                 int32_t pixel_type, char* err, size_t err_len) {
    try {
        return exr_save_impl(path, rgba, width, height, pixel_type, err, err_len);
    } catch (const std::bad_alloc&) {
        return fail(err, err_len, LT_ERR_OUT_OF_MEMORY, "lt_exr_save: out of host memory");
    }
}

// Limits on what a file may ask of the reader: channels per file and bytes of
// one decoded chunk (ZIP: 16 scanlines); beyond them the file is rejected.
constexpr size_t kMaxChannels = 64;
constexpr uint64_t kMaxChunkBytes = 1ull << 31;

static lt_status exr_load_impl(const char* path, float* rgba, size_t capacity, uint32_t* width_out,
                               uint32_t* height_out, char* err, size_t err_len) {
    if (!path || !width_out || !height_out)
        return fail(err, err_len, LT_ERR_INVALID_ARG, "lt_exr_load: null pointer");
    FILE* fp = fopen(path, "rb");
    if (!fp) return fail(err, err_len, LT_ERR_INVALID_ARG, "lt_exr_load: cannot open %s", path);
    std::vector<uint8_t> file;
    uint8_t chunk[1 << 16];
    size_t got;
    while ((got = fread(chunk, 1, sizeof(chunk), fp)) > 0) file.insert(file.end(), chunk, chunk + got);
    fclose(fp);
    Reader r{file.data(), file.size()};
    if (r.u32() != kMagic || !r.ok) return fail(err, err_len, LT_ERR_PARSE, "%s: not an OpenEXR file", path);
    const uint32_t version = r.u32();
    if ((version & 0xffu) != 2u || (version & 0xffffff00u) != 0u)
        return fail(err, err_len, LT_ERR_UNSUPPORTED, "%s: version word 0x%x (only single-part scanline files)",
                    path, version);
    std::vector<Channel> chans;
    int32_t box[4] = {0, 0, -1, -1};
    bool have_box = false;
    uint8_t comp = 255;
    while (r.ok && r.need(1) && r.p[r.at] != 0) {
        const std::string name = r.cstr(), type = r.cstr();
        const uint32_t size = r.u32();
        if (!r.ok || !r.need(size)) break;
        Reader a{r.p + r.at, size};
        if (name == "channels" && type == "chlist") {
            while (a.ok && a.need(1) && a.p[a.at] != 0) {
                Channel c;
                c.name = a.cstr();
                c.type = (int32_t)a.u32();
                a.u32();  // pLinear + reserved
                const uint32_t xs = a.u32(), ys = a.u32();
                if (xs != 1u || ys != 1u)
                    return fail(err, err_len, LT_ERR_UNSUPPORTED, "%s: subsampled channel %s", path, c.name.c_str());
                if (c.type < kUint || c.type > kFloat)
                    return fail(err, err_len, LT_ERR_PARSE, "%s: channel %s has pixel type %d", path,
                                c.name.c_str(), c.type);
                c.slot = c.name == "R" ? 0 : c.name == "G" ? 1 : c.name == "B" ? 2 : c.name == "A" ? 3 : -1;
                if (chans.size() == kMaxChannels)
                    return fail(err, err_len, LT_ERR_UNSUPPORTED, "%s: more than %zu channels", path, kMaxChannels);
                chans.push_back(c);
            }
            if (!a.ok) return fail(err, err_len, LT_ERR_PARSE, "%s: truncated channel list", path);
        } else if (name == "compression" && size == 1) {
            comp = a.p[0];
        } else if (name == "dataWindow" && size == 16) {
            for (int i = 0; i < 4; i++) box[i] = (int32_t)a.u32();
            have_box = true;
        }
        r.at += size;
    }
    if (!r.ok || !r.need(1)) return fail(err, err_len, LT_ERR_PARSE, "%s: truncated header", path);
    r.at++;  // header terminator
    if (chans.empty() || !have_box) return fail(err, err_len, LT_ERR_PARSE, "%s: no channels or no dataWindow", path);
    if (comp != kNone && comp != kZips && comp != kZip)
        return fail(err, err_len, LT_ERR_UNSUPPORTED, "%s: compression %u (NONE, ZIPS and ZIP are read)", path, comp);
    const int64_t w64 = (int64_t)box[2] - box[0] + 1, h64 = (int64_t)box[3] - box[1] + 1;
    if (w64 <= 0 || h64 <= 0 || w64 > (1 << 20) || h64 > (1 << 20) || w64 * h64 > (1ll << 30))
        return fail(err, err_len, LT_ERR_PARSE, "%s: bad dataWindow", path);
    const uint32_t W = (uint32_t)w64, H = (uint32_t)h64;
    *width_out = W;
    *height_out = H;
    if (!rgba) return LT_OK;  // size query
    if (capacity < 4ull * W * H)
        return fail(err, err_len, LT_ERR_INVALID_ARG, "lt_exr_load: %s needs %llu floats, capacity %zu", path,
                    (unsigned long long)(4ull * W * H), capacity);
    uint64_t line_bytes = 0;
    for (const Channel& c : chans) line_bytes += (uint64_t)type_bytes(c.type) * W;
    const uint32_t lines_per_chunk = comp == kZip ? 16u : 1u;
    if (line_bytes * lines_per_chunk > kMaxChunkBytes)
        return fail(err, err_len, LT_ERR_UNSUPPORTED, "%s: a chunk of %llu bytes", path,
                    (unsigned long long)(line_bytes * lines_per_chunk));
    const uint32_t nchunks = (H + lines_per_chunk - 1) / lines_per_chunk;
    for (size_t i = 0; i < 4ull * W * H; i++) rgba[i] = (i & 3) == 3 ? 1.f : 0.f;
    std::vector<uint8_t> buf, tmp;
    std::vector<uint8_t> seen(H, 0);  // every scanline exactly once
    for (uint32_t k = 0; k < nchunks; k++) {
        Reader t{file.data(), file.size(), r.at + 8ull * k};
        const uint64_t off = t.u64();
        if (!t.ok || off > file.size()) return fail(err, err_len, LT_ERR_PARSE, "%s: bad offset table", path);
        Reader c{file.data(), file.size(), (size_t)off};
        const int32_t y0 = (int32_t)c.u32();
        const uint32_t bytes = c.u32();
        if (!c.ok || !c.need(bytes) || y0 < box[1] || y0 > box[3] || (int64_t)(y0 - box[1]) % lines_per_chunk)
            return fail(err, err_len, LT_ERR_PARSE, "%s: bad chunk %u", path, k);
        const uint32_t row0 = (uint32_t)(y0 - box[1]);
        const uint32_t nl = std::min(lines_per_chunk, H - row0);
        for (uint32_t l = 0; l < nl; l++) {
            if (seen[row0 + l]) return fail(err, err_len, LT_ERR_PARSE, "%s: scanline %u given twice", path, row0 + l);
            seen[row0 + l] = 1;
        }
        const uint64_t expect = line_bytes * nl;
        const uint8_t* src = c.p + c.at;
        if (bytes != expect) {  // compressed chunk (an incompressible one is stored raw)
            if (comp == kNone) return fail(err, err_len, LT_ERR_PARSE, "%s: chunk %u has %u bytes", path, k, bytes);
            buf.resize(expect);
            uLongf dst = (uLongf)expect;
            if (uncompress(buf.data(), &dst, src, bytes) != Z_OK || dst != expect)
                return fail(err, err_len, LT_ERR_PARSE, "%s: chunk %u does not inflate", path, k);
            unfilter(buf, tmp);
            src = buf.data();
        }
        for (uint32_t l = 0; l < nl; l++) {
            float* row = rgba + 4ull * W * (row0 + l);
            for (const Channel& ch : chans) {
                const uint32_t bpc = type_bytes(ch.type);
                if (ch.slot >= 0)
                    for (uint32_t x = 0; x < W; x++) {
                        const uint8_t* q = src + (size_t)x * bpc;
                        float v;
                        if (ch.type == kHalf) {
                            v = f16_to_f32((uint16_t)(q[0] | q[1] << 8));
                        } else {
                            const uint32_t u = (uint32_t)q[0] | (uint32_t)q[1] << 8 | (uint32_t)q[2] << 16 |
                                               (uint32_t)q[3] << 24;
                            if (ch.type == kFloat)
                                memcpy(&v, &u, 4);
                            else
                                v = (float)u;
                        }
                        row[4 * x + ch.slot] = v;
                    }
                src += (size_t)bpc * W;
            }
        }
    }
    for (uint32_t y = 0; y < H; y++)
        if (!seen[y]) return fail(err, err_len, LT_ERR_PARSE, "%s: scanline %u missing", path, y);
    return LT_OK;
}

extern "C" lt_status lt_exr_load(const char* path, float* rgba, size_t capacity, uint32_t* width_out,
                                 uint32_t* height_out, char* err, size_t err_len) {
    try {
        return exr_load_impl(path, rgba, capacity, width_out, height_out, err, err_len);
    } catch (const std::bad_alloc&) {
        return fail(err, err_len, LT_ERR_OUT_OF_MEMORY, "%s: out of host memory", path ? path : "lt_exr_load");
    } catch (...) {
        return fail(err, err_len, LT_ERR_PARSE, "%s: unreadable file", path ? path : "lt_exr_load");
    }
}
