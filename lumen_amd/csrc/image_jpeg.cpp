// image_jpeg.cpp — baseline JPEG decoder for scene textures.
//
// The reference decodes a material's "texture" file with stb_image into
// 8-bit RGBA (src/RayTracer/LumenScene.cpp:205-212, stbi_load(..., 4)) and
// samples it as VK_FORMAT_R8G8B8A8_SRGB.  This decoder follows ITU-T T.81
// for the sequential Huffman processes (SOF0 baseline, SOF1 extended with
// 8-bit samples): DQT, DHT, DRI and restart markers, interleaved and
// non-interleaved scans, any sampling factors up to 4, one (grey) or three
// (YCbCr, or RGB when an Adobe marker says so or the component ids are
// 'R','G','B') components.  Progressive, arithmetic-coded, 12-bit and
// lossless files are refused with LT_ERR_UNSUPPORTED.
//
// Reconstruction: dequantised coefficients through a separable float IDCT
// (T.81 A.3.3) with +128 level shift, rounded and clamped; chroma planes of
// half resolution upsampled with the triangle (3/4, 1/4) filter in each
// halved direction, other ratios replicated; YCbCr to RGB by the JFIF
// equations.  These are the steps stb_image takes; its integer IDCT and
// fixed-point colour conversion round differently, so decoded values agree
// with it to within a few units of the last 8-bit place, not bit for bit.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lt.h"
#include "scene.h"

namespace lt {
namespace {

// zig-zag position k -> natural (row-major) index in the 8x8 block
constexpr uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huff {
    bool defined = false;
    // canonical decoding (T.81 F.2.2.3): per code length l = 1..16 the
    // largest code of that length (-1: none) and where its values start
    int32_t maxcode[18];
    int32_t valptr[17];
    int32_t mincode[17];
    uint8_t vals[256];
};

struct Comp {
    int id = 0, h = 1, v = 1, tq = 0;
    int td = 0, ta = 0;   // DC / AC table of the current scan
    int bw = 0, bh = 0;   // blocks per line / column (padded to whole MCUs)
    int dc_pred = 0;
    std::vector<int16_t> coef;  // bw * bh blocks of 64 dequantised coefficients (natural order)
};

struct Decoder {
    const uint8_t* p;
    const uint8_t* end;
    std::string err;
    uint16_t qt[4][64];  // natural order
    bool qt_defined[4] = {false, false, false, false};
    Huff dc[4], ac[4];
    std::vector<Comp> comps;
    int width = 0, height = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
    int restart_interval = 0;
    bool adobe = false;
    int adobe_transform = -1;
    bool jfif = false;
    bool frame_seen = false;
    // entropy-coded segment bit reader
    uint32_t bitbuf = 0;
    int bitcnt = 0;
    bool marker_hit = false;

    bool fail(const char* m) {
        if (err.empty()) err = m;
        return false;
    }
    int u8() { return p < end ? *p++ : (marker_hit = true, 0); }
    int u16() {
        int a = u8();
        return (a << 8) | u8();
    }

    // fills bits, removing stuffed zero bytes; at a marker it feeds zeros
    // (T.81 F.2.2.5: decoding must not read past the marker)
    void fill() {
        while (bitcnt <= 24) {
            int b = 0;
            if (!marker_hit && p < end) {
                if (p[0] == 0xFF) {
                    if (p + 1 < end && p[1] == 0x00) {
                        b = 0xFF;
                        p += 2;
                    } else {
                        marker_hit = true;
                    }
                } else {
                    b = *p++;
                }
            }
            bitbuf |= (uint32_t)b << (24 - bitcnt);
            bitcnt += 8;
        }
    }
    int bits(int n) {
        if (n == 0) return 0;
        fill();
        const int v = (int)(bitbuf >> (32 - n));
        bitbuf <<= n;
        bitcnt -= n;
        return v;
    }
    // T.81 F.2.2.1 EXTEND
    static int extend(int v, int t) { return v < (1 << (t - 1)) ? v - (1 << t) + 1 : v; }
    int decode(const Huff& h) {
        fill();
        int code = 0;
        for (int l = 1; l <= 16; ++l) {
            code = (code << 1) | (int)(bitbuf >> 31);
            bitbuf <<= 1;
            --bitcnt;
            if (h.maxcode[l] >= 0 && code <= h.maxcode[l]) return h.vals[h.valptr[l] + code - h.mincode[l]];
        }
        return -1;  // no code of 16 bits or less: corrupt data
    }
    void reset_bits() {
        bitbuf = 0;
        bitcnt = 0;
    }

    bool read_dqt(int len) {
        const uint8_t* stop = p + len;
        while (p < stop) {
            const int pq_tq = u8();
            const int pq = pq_tq >> 4, tq = pq_tq & 15;
            if (tq > 3 || pq > 1) return fail("bad DQT");
            for (int k = 0; k < 64; ++k) qt[tq][kZigzag[k]] = (uint16_t)(pq ? u16() : u8());
            qt_defined[tq] = true;
        }
        return p == stop || fail("bad DQT length");
    }
    bool read_dht(int len) {
        const uint8_t* stop = p + len;
        while (p < stop) {
            const int tc_th = u8();
            const int tc = tc_th >> 4, th = tc_th & 15;
            if (tc > 1 || th > 3) return fail("bad DHT");
            Huff& h = tc ? ac[th] : dc[th];
            int counts[17], total = 0;
            for (int l = 1; l <= 16; ++l) total += counts[l] = u8();
            if (total > 256) return fail("bad DHT counts");
            for (int i = 0; i < total; ++i) h.vals[i] = (uint8_t)u8();
            // canonical codes (T.81 C.2, F.2.2.3)
            int code = 0, k = 0;
            for (int l = 1; l <= 16; ++l) {
                h.valptr[l] = k;
                h.mincode[l] = code;
                code += counts[l];
                k += counts[l];
                h.maxcode[l] = counts[l] ? code - 1 : -1;
                if (counts[l] && code - 1 >= (1 << l)) return fail("bad DHT code space");
                code <<= 1;
            }
            h.maxcode[17] = 0x7fffffff;
            h.defined = true;
        }
        return p == stop || fail("bad DHT length");
    }
    bool read_sof(int len, int type) {
        if (frame_seen) return fail("more than one frame");
        frame_seen = true;
        if (type != 0xC0 && type != 0xC1) return fail("unsupported JPEG process");
        if (u8() != 8) return fail("unsupported sample precision");
        height = u16();
        width = u16();
        const int nc = u8();
        if (width <= 0 || height <= 0) return fail("unsupported image size (DNL)");
        if (width > 16384 || height > 16384) return fail("image too large");
        if (nc != 1 && nc != 3) return fail("unsupported component count");
        if (len != 6 + 3 * nc) return fail("bad SOF length");
        comps.resize(nc);
        for (auto& c : comps) {
            c.id = u8();
            const int hv = u8();
            c.h = hv >> 4;
            c.v = hv & 15;
            c.tq = u8();
            if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) return fail("bad SOF component");
            hmax = std::max(hmax, c.h);
            vmax = std::max(vmax, c.v);
        }
        mcux = (width + 8 * hmax - 1) / (8 * hmax);
        mcuy = (height + 8 * vmax - 1) / (8 * vmax);
        for (auto& c : comps) {
            c.bw = mcux * c.h;
            c.bh = mcuy * c.v;
            c.coef.assign((size_t)c.bw * c.bh * 64, 0);
        }
        return true;
    }

    bool decode_block(Comp& c, int bx, int by) {
        const Huff& hd = dc[c.td];
        const Huff& ha = ac[c.ta];
        int16_t* out = &c.coef[((size_t)by * c.bw + bx) * 64];
        const uint16_t* q = qt[c.tq];
        const int t = decode(hd);
        if (t < 0 || t > 11) return fail("corrupt DC code");
        c.dc_pred += t ? extend(bits(t), t) : 0;
        int32_t v = c.dc_pred * (int32_t)q[0];
        out[0] = (int16_t)std::max(-32768, std::min(32767, v));
        for (int k = 1; k < 64;) {
            const int rs = decode(ha);
            if (rs < 0) return fail("corrupt AC code");
            const int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r != 15) break;  // EOB
                k += 16;             // ZRL
                continue;
            }
            k += r;
            if (k > 63) return fail("AC index out of range");
            v = extend(bits(s), s) * (int32_t)q[kZigzag[k]];
            out[kZigzag[k]] = (int16_t)std::max(-32768, std::min(32767, v));
            ++k;
        }
        return true;
    }

    // at a restart marker: T.81 F.2.1.3.1 (predictions back to zero)
    bool restart() {
        reset_bits();
        // skip to the marker
        while (p + 1 < end && !(p[0] == 0xFF && p[1] >= 0xD0 && p[1] <= 0xD7)) ++p;
        if (p + 1 >= end) return fail("missing restart marker");
        p += 2;
        marker_hit = false;
        for (auto& c : comps) c.dc_pred = 0;
        return true;
    }

    bool read_sos(int len) {
        if (!frame_seen) return fail("scan before frame");
        const int ns = u8();
        if (ns < 1 || ns > (int)comps.size() || len != 4 + 2 * ns) return fail("bad SOS");
        std::vector<Comp*> sc;
        for (int i = 0; i < ns; ++i) {
            const int id = u8(), tt = u8();
            Comp* found = nullptr;
            for (auto& c : comps)
                if (c.id == id) found = &c;
            if (!found) return fail("scan names an unknown component");
            found->td = tt >> 4;
            found->ta = tt & 15;
            if (found->td > 3 || found->ta > 3 || !dc[found->td].defined || !ac[found->ta].defined)
                return fail("scan uses an undefined Huffman table");
            if (!qt_defined[found->tq]) return fail("component uses an undefined quantisation table");
            sc.push_back(found);
        }
        const int ss = u8(), se = u8(), ahal = u8();
        if (ss != 0 || se != 63 || ahal != 0) return fail("unsupported spectral selection");
        reset_bits();
        marker_hit = false;
        for (auto* c : sc) c->dc_pred = 0;
        int todo = restart_interval;
        auto step = [&]() -> bool {
            if (restart_interval && todo-- == 0) {
                if (!restart()) return false;
                todo = restart_interval - 1;
            }
            return true;
        };
        if (ns == 1) {
            // non-interleaved: blocks in raster order over the component's own
            // (unpadded) extent, T.81 A.2.2
            Comp& c = *sc[0];
            const int cw = (width * c.h + hmax - 1) / hmax, ch = (height * c.v + vmax - 1) / vmax;
            const int nbx = (cw + 7) / 8, nby = (ch + 7) / 8;
            for (int by = 0; by < nby; ++by)
                for (int bx = 0; bx < nbx; ++bx) {
                    if (!step() || !decode_block(c, bx, by)) return false;
                }
        } else {
            for (int my = 0; my < mcuy; ++my)
                for (int mx = 0; mx < mcux; ++mx) {
                    if (!step()) return false;
                    for (auto* c : sc)
                        for (int y = 0; y < c->v; ++y)
                            for (int x = 0; x < c->h; ++x)
                                if (!decode_block(*c, mx * c->h + x, my * c->v + y)) return false;
                }
        }
        // leave p at the next marker
        reset_bits();
        while (p + 1 < end && !(p[0] == 0xFF && p[1] != 0x00 && !(p[1] >= 0xD0 && p[1] <= 0xD7))) ++p;
        marker_hit = false;
        return true;
    }

    bool parse() {
        if (end - p < 4 || p[0] != 0xFF || p[1] != 0xD8) return fail("not a JPEG file");
        p += 2;
        bool scanned = false;
        while (p < end) {
            if (*p != 0xFF) return fail("marker expected");
            while (p < end && *p == 0xFF) ++p;  // fill bytes
            if (p >= end) break;
            const int m = *p++;
            if (m == 0xD9) break;  // EOI
            if (m >= 0xD0 && m <= 0xD7) continue;
            if (end - p < 2) return fail("truncated marker");
            const int len = u16() - 2;
            if (len < 0 || len > end - p) return fail("truncated segment");
            const uint8_t* next = p + len;
            switch (m) {
                case 0xDB:
                    if (!read_dqt(len)) return false;
                    break;
                case 0xC4:
                    if (!read_dht(len)) return false;
                    break;
                case 0xDD:
                    restart_interval = u16();
                    break;
                case 0xC0:
                case 0xC1:
                    if (!read_sof(len, m)) return false;
                    break;
                case 0xC2:
                case 0xC3:
                case 0xC5:
                case 0xC6:
                case 0xC7:
                case 0xC9:
                case 0xCA:
                case 0xCB:
                case 0xCD:
                case 0xCE:
                case 0xCF:
                    return fail("unsupported JPEG process (progressive, lossless or arithmetic)");
                case 0xDA:
                    if (!read_sos(len)) return false;
                    scanned = true;
                    continue;  // p is already at the next marker
                case 0xE0:
                    jfif = len >= 5 && !memcmp(p, "JFIF", 5);
                    break;
                case 0xEE:
                    if (len >= 12 && !memcmp(p, "Adobe", 5)) {
                        adobe = true;
                        adobe_transform = p[11];
                    }
                    break;
                default:
                    break;  // APPn, COM and others
            }
            p = next;
        }
        return scanned || fail("no scan");
    }
};

// 1-D 8-point IDCT basis: c[x][u] = C(u)/2 cos((2x+1) u pi / 16), C(0) = 1/sqrt 2
struct IdctBasis {
    float c[8][8];
    IdctBasis() {
        for (int x = 0; x < 8; ++x)
            for (int u = 0; u < 8; ++u)
                c[x][u] = (float)((u == 0 ? std::sqrt(0.5) : 1.0) * 0.5 * std::cos((2 * x + 1) * u * M_PI / 16.0));
    }
};

void idct_block(const int16_t* in, uint8_t* out, int stride) {
    static const IdctBasis B;
    float tmp[64];
    for (int v = 0; v < 8; ++v)  // rows: x direction
        for (int x = 0; x < 8; ++x) {
            float s = 0.0f;
            for (int u = 0; u < 8; ++u) s += B.c[x][u] * (float)in[v * 8 + u];
            tmp[v * 8 + x] = s;
        }
    for (int x = 0; x < 8; ++x)
        for (int y = 0; y < 8; ++y) {
            float s = 0.0f;
            for (int v = 0; v < 8; ++v) s += B.c[y][v] * tmp[v * 8 + x];
            const int r = (int)std::lround(s + 128.0f);
            out[y * stride + x] = (uint8_t)std::max(0, std::min(255, r));
        }
}

uint8_t clamp8(int v) { return (uint8_t)std::max(0, std::min(255, v)); }

}  // namespace

// Decodes a JPEG byte stream into 8-bit RGBA (alpha 255), rows top to bottom.
lt_status decode_jpeg(const uint8_t* data, size_t n, std::vector<uint8_t>& rgba, uint32_t& w, uint32_t& h,
                      std::string& err) {
    Decoder d;
    d.p = data;
    d.end = data + n;
    if (!d.parse()) {
        err = d.err;
        return d.err.find("unsupported") == 0 ? LT_ERR_UNSUPPORTED : LT_ERR_PARSE;
    }
    w = (uint32_t)d.width;
    h = (uint32_t)d.height;
    // component planes at their own resolution (padded)
    std::vector<std::vector<uint8_t>> planes(d.comps.size());
    for (size_t i = 0; i < d.comps.size(); ++i) {
        Comp& c = d.comps[i];
        const int pw = c.bw * 8;
        planes[i].resize((size_t)pw * c.bh * 8);
        for (int by = 0; by < c.bh; ++by)
            for (int bx = 0; bx < c.bw; ++bx)
                idct_block(&c.coef[((size_t)by * c.bw + bx) * 64], &planes[i][(size_t)by * 8 * pw + bx * 8], pw);
    }
    // upsample every plane to full resolution
    const int W = d.width, H = d.height;
    std::vector<std::vector<uint8_t>> full(d.comps.size(), std::vector<uint8_t>((size_t)W * H));
    for (size_t i = 0; i < d.comps.size(); ++i) {
        const Comp& c = d.comps[i];
        const int pw = c.bw * 8;
        const int fx = d.hmax / c.h, fy = d.vmax / c.v;
        const bool exact = d.hmax % c.h == 0 && d.vmax % c.v == 0;
        const int cw = (W * c.h + d.hmax - 1) / d.hmax, ch = (H * c.v + d.vmax - 1) / d.vmax;
        const uint8_t* src = planes[i].data();
        uint8_t* dst = full[i].data();
        auto at = [&](int x, int y) { return (int)src[(size_t)y * pw + x]; };
        for (int y = 0; y < H; ++y) {
            // vertical: near row y / fy; for fy == 2 the far row is the one
            // above (even y) or below (odd y), clamped to the plane
            const int sy = exact ? y / fy : y * c.v / d.vmax;
            int far_y = sy;
            if (exact && fy == 2) far_y = std::max(0, std::min(ch - 1, (y & 1) ? sy + 1 : sy - 1));
            for (int x = 0; x < W; ++x) {
                const int sx = exact ? x / fx : x * c.h / d.hmax;
                int val;
                if (exact && fy == 2 && fx == 2) {
                    // 3/4 near + 1/4 far in each direction
                    const int far_x = std::max(0, std::min(cw - 1, (x & 1) ? sx + 1 : sx - 1));
                    const int t_near = 3 * at(sx, sy) + at(sx, far_y);
                    const int t_far = 3 * at(far_x, sy) + at(far_x, far_y);
                    val = (3 * t_near + t_far + 8) >> 4;
                } else if (exact && fx == 2 && fy == 1) {
                    const int far_x = std::max(0, std::min(cw - 1, (x & 1) ? sx + 1 : sx - 1));
                    val = (3 * at(sx, sy) + at(far_x, sy) + 2) >> 2;
                } else if (exact && fx == 1 && fy == 2) {
                    val = (3 * at(sx, sy) + at(sx, far_y) + 2) >> 2;
                } else {
                    val = at(sx, sy);
                }
                dst[(size_t)y * W + x] = (uint8_t)val;
            }
        }
    }
    rgba.assign((size_t)W * H * 4, 255);
    const bool rgb_ids = d.comps.size() == 3 && d.comps[0].id == 'R' && d.comps[1].id == 'G' && d.comps[2].id == 'B';
    const bool no_transform = d.comps.size() == 3 && (rgb_ids || (d.adobe && d.adobe_transform == 0 && !d.jfif));
    for (size_t k = 0; k < (size_t)W * H; ++k) {
        uint8_t* o = &rgba[4 * k];
        if (d.comps.size() == 1) {
            o[0] = o[1] = o[2] = full[0][k];
        } else if (no_transform) {
            o[0] = full[0][k];
            o[1] = full[1][k];
            o[2] = full[2][k];
        } else {
            // JFIF YCbCr -> RGB
            const float Y = full[0][k], cb = (float)full[1][k] - 128.0f, cr = (float)full[2][k] - 128.0f;
            o[0] = clamp8((int)std::lround(Y + 1.402f * cr));
            o[1] = clamp8((int)std::lround(Y - 0.344136f * cb - 0.714136f * cr));
            o[2] = clamp8((int)std::lround(Y + 1.772f * cb));
        }
    }
    return LT_OK;
}

}  // namespace lt

extern "C" lt_status lt_image_load_rgba8(const char* path, uint8_t* rgba, size_t capacity, uint32_t* width,
                                         uint32_t* height, char* err, size_t err_len) {
    auto fail = [&](lt_status st, const std::string& m) {
        if (err && err_len) snprintf(err, err_len, "%s", m.c_str());
        return st;
    };
    if (!path || !width || !height) return fail(LT_ERR_INVALID_ARG, "lt_image_load_rgba8: null pointer");
    FILE* fp = fopen(path, "rb");
    if (!fp) return fail(LT_ERR_INVALID_ARG, std::string("lt_image_load_rgba8: cannot open ") + path);
    std::vector<uint8_t> file;
    uint8_t chunk[1 << 16];
    size_t got;
    while ((got = fread(chunk, 1, sizeof(chunk), fp)) > 0) file.insert(file.end(), chunk, chunk + got);
    fclose(fp);
    std::vector<uint8_t> out;
    uint32_t w = 0, h = 0;
    std::string e;
    const lt_status st = lt::decode_jpeg(file.data(), file.size(), out, w, h, e);
    if (st != LT_OK) return fail(st, std::string(path) + ": " + e);
    *width = w;
    *height = h;
    if (!rgba) return LT_OK;
    if (capacity < out.size()) return fail(LT_ERR_INVALID_ARG, "lt_image_load_rgba8: capacity below 4 * width * height");
    memcpy(rgba, out.data(), out.size());
    return LT_OK;
}
