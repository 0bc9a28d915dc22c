// Bitmap loading for textures and environment maps: util::BitmapTexture::Load
// (util/texture.cpp:87-174), which reads EXR through tinyexr and everything
// else through stb_image.  Written here from the file-format specifications:
//   OpenEXR  scanline files, compression NONE / RLE / ZIPS / ZIP, pixel types
//            UINT / HALF / FLOAT (channels R, G, B, A or Y);
//   Radiance RGBE (.hdr / .pic), flat and new run-length scanlines;
//   PNG      8- and 16-bit (and 1/2/4-bit gray / palette), plain or Adam7-interlaced;
//   BMP      uncompressed palette / 16 / 24 / 32-bit, bit fields;
//   TGA      color-mapped, true-color and gray, raw or run-length encoded;
//   PNM      binary P5 / P6;
//   JPEG     baseline, extended sequential and progressive Huffman (pt_jpeg.cpp);
//   GIF      the first image (87a / 89a), transparency, interlacing;
//   PSD      the composite image, 8/16-bit RGB, raw or PackBits.
// Output: RGBA float, row 0 = the first (top) row of the file, as both
// loaders of the reference return it.  8-bit data is linearized with
// (v / 255)^2.2 as in util/texture.cpp:106-111 (16-bit PNG channels keep their
// high byte first, as stb_image's 8-bit path does); HDR and EXR stay linear.
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pt_scene.h"

namespace pupil {

// One size limit for every reader, applied to the header's dimensions before
// any pixel buffer is allocated (as stb_image limits each side): at most
// 2^24 pixels per side and 2^28 pixels in all (4 GiB of float RGBA).
bool image_size_ok(uint64_t W, uint64_t H, std::string &err) {
    if (W == 0 || H == 0 || W > kMaxImageSide || H > kMaxImageSide || W * H > kMaxImagePixels) {
        err = "unsupported: image of " + std::to_string(W) + "x" + std::to_string(H) +
              " pixels (at most 2^24 per side and 2^28 in all)";
        return false;
    }
    return true;
}

namespace {

bool read_file(const std::string &path, std::vector<unsigned char> &out, std::string &err) {
    FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) {
        err = "io: cannot open image " + path;
        return false;
    }
    std::fseek(fp, 0, SEEK_END);
    long n = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    if (n < 0) {
        std::fclose(fp);
        err = "io: cannot size image " + path;
        return false;
    }
    out.resize((size_t)n);
    size_t rd = n ? std::fread(out.data(), 1, (size_t)n, fp) : 0;
    std::fclose(fp);
    if (rd != (size_t)n) {
        err = "io: short read " + path;
        return false;
    }
    return true;
}

std::string lower_ext(const std::string &path) {
    size_t dot = path.find_last_of('.');
    size_t slash = path.find_last_of("/\\");
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return "";
    std::string e = path.substr(dot + 1);
    for (char &c : e) c = (char)std::tolower((unsigned char)c);
    return e;
}

float half_to_float(uint16_t h) {
    const uint32_t s = (uint32_t)(h >> 15) << 31;
    const uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) {
            bits = s;
        } else {  // subnormal: m * 2^-24
            float f = (float)m * 5.9604644775390625e-8f;
            std::memcpy(&bits, &f, 4);
            bits |= s;
        }
    } else if (e == 31) {
        bits = s | 0x7F800000u | (m << 13);
    } else {
        bits = s | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

// ------------------------------------------------------------------ OpenEXR
struct ExrChannel {
    std::string name;
    int32_t type;  // 0 uint, 1 half, 2 float
    int32_t xs, ys;
};

bool exr_undo_predictor_and_reorder(std::vector<unsigned char> &buf) {
    for (size_t i = 1; i < buf.size(); ++i) buf[i] = (unsigned char)(buf[i - 1] + buf[i] - 128);
    std::vector<unsigned char> out(buf.size());
    const size_t half = (buf.size() + 1) / 2;
    for (size_t i = 0, a = 0, b = half; i < buf.size(); ++i) out[i] = (i & 1) ? buf[b++] : buf[a++];
    buf.swap(out);
    return true;
}

bool exr_rle_decode(const unsigned char *in, size_t n, size_t expect, std::vector<unsigned char> &out) {
    out.clear();
    size_t i = 0;
    while (i < n) {
        const int8_t c = (int8_t)in[i++];
        if (c < 0) {
            const size_t run = (size_t)(-(int)c);
            if (i + run > n) return false;
            out.insert(out.end(), in + i, in + i + run);
            i += run;
        } else {
            if (i >= n) return false;
            out.insert(out.end(), (size_t)c + 1, in[i++]);
        }
        if (out.size() > expect) return false;
    }
    return out.size() == expect;
}

bool load_exr(const std::vector<unsigned char> &f, uint32_t &w, uint32_t &h, std::vector<float> &rgba,
              std::string &err) {
    size_t p = 0;
    auto need = [&](size_t k) { return p + k <= f.size(); };
    auto rd_i32 = [&](int32_t &v) {
        if (!need(4)) return false;
        std::memcpy(&v, &f[p], 4);
        p += 4;
        return true;
    };
    auto rd_str = [&](std::string &s) {
        size_t e = p;
        while (e < f.size() && f[e] != 0) ++e;
        if (e >= f.size()) return false;
        s.assign(reinterpret_cast<const char *>(&f[p]), e - p);
        p = e + 1;
        return true;
    };
    int32_t magic = 0, version = 0;
    if (!rd_i32(magic) || magic != 20000630 || !rd_i32(version)) {
        err = "parse: not an OpenEXR file";
        return false;
    }
    // version flags: 0x200 tiled, 0x800 deep data, 0x1000 multipart
    if ((version & 0xFF) != 2 || (version & 0x200) || (version & 0x800) || (version & 0x1000)) {
        err = "unsupported: EXR files other than single-part scanline images";
        return false;
    }
    std::vector<ExrChannel> chans;
    int compression = -1, line_order = 0;
    int32_t dw[4] = {0, 0, -1, -1};
    for (;;) {
        std::string name, type;
        if (!rd_str(name)) {
            err = "parse: truncated EXR header";
            return false;
        }
        if (name.empty()) break;
        int32_t size = 0;
        if (!rd_str(type) || !rd_i32(size) || size < 0 || !need((size_t)size)) {
            err = "parse: truncated EXR header";
            return false;
        }
        const size_t end = p + (size_t)size;
        if (name == "channels" && type == "chlist") {
            while (p < end && f[p] != 0) {
                ExrChannel c;
                int32_t dummy;
                if (!rd_str(c.name) || !rd_i32(c.type) || !rd_i32(dummy) || !rd_i32(c.xs) || !rd_i32(c.ys)) {
                    err = "parse: bad EXR channel list";
                    return false;
                }
                chans.push_back(c);
            }
        } else if (name == "compression") {
            compression = f[p];
        } else if (name == "dataWindow" && size == 16) {
            std::memcpy(dw, &f[p], 16);
        } else if (name == "lineOrder") {
            line_order = f[p];
        }
        p = end;
    }
    (void)line_order;  // chunks carry their own y; the order does not matter here
    if (dw[2] < dw[0] || dw[3] < dw[1] || chans.empty()) {
        err = "parse: EXR without data window or channels";
        return false;
    }
    const int64_t W = (int64_t)dw[2] - dw[0] + 1, H = (int64_t)dw[3] - dw[1] + 1;
    if (W > 65536 || H > 65536) {
        err = "unsupported: EXR image larger than 65536 pixels per side";
        return false;
    }
    if (!image_size_ok((uint64_t)W, (uint64_t)H, err)) return false;
    int lines = 0;
    switch (compression) {
        case 0: case 1: case 2: lines = 1; break;
        case 3: lines = 16; break;
        default:
            err = "unsupported: EXR compression " + std::to_string(compression) + " (NONE, RLE, ZIPS and ZIP are read)";
            return false;
    }
    size_t pix_bytes = 0;
    for (auto &c : chans) {
        if (c.xs != 1 || c.ys != 1 || c.type < 0 || c.type > 2) {
            err = "unsupported: subsampled or unknown EXR channel " + c.name;
            return false;
        }
        pix_bytes += c.type == 1 ? 2 : 4;
    }
    int ci[4] = {-1, -1, -1, -1};  // R G B A
    int cy = -1;
    for (size_t i = 0; i < chans.size(); ++i) {
        const std::string &n = chans[i].name;
        if (n == "R") ci[0] = (int)i;
        else if (n == "G") ci[1] = (int)i;
        else if (n == "B") ci[2] = (int)i;
        else if (n == "A") ci[3] = (int)i;
        else if (n == "Y") cy = (int)i;
    }
    if ((ci[0] < 0 || ci[1] < 0 || ci[2] < 0) && cy < 0) {
        err = "unsupported: EXR without R, G, B (or Y) channels";
        return false;
    }
    w = (uint32_t)W;
    h = (uint32_t)H;
    rgba.assign((size_t)W * H * 4, 0.f);
    for (size_t i = 0; i < (size_t)W * H; ++i) rgba[4 * i + 3] = 1.f;
    const int64_t chunks = (H + lines - 1) / lines;
    if (!need((size_t)chunks * 8)) {
        err = "parse: truncated EXR offset table";
        return false;
    }
    std::vector<uint64_t> offs((size_t)chunks);
    std::memcpy(offs.data(), &f[p], (size_t)chunks * 8);
    std::vector<unsigned char> buf;
    for (int64_t k = 0; k < chunks; ++k) {
        p = (size_t)offs[(size_t)k];
        int32_t y0 = 0, size = 0;
        if (offs[(size_t)k] >= f.size() || !rd_i32(y0) || !rd_i32(size) || size < 0 || !need((size_t)size)) {
            err = "parse: bad EXR chunk";
            return false;
        }
        const int64_t ys = (int64_t)y0 - dw[1];
        if (ys < 0 || ys >= H) {
            err = "parse: EXR chunk outside the data window";
            return false;
        }
        const int64_t ny = std::min<int64_t>(lines, H - ys);
        const size_t expect = (size_t)ny * (size_t)W * pix_bytes;
        const unsigned char *src = &f[p];
        if ((size_t)size == expect) {  // stored uncompressed (always for NONE)
            buf.assign(src, src + expect);
        } else if (compression == 1) {
            if (!exr_rle_decode(src, (size_t)size, expect, buf)) {
                err = "parse: bad EXR RLE chunk";
                return false;
            }
            exr_undo_predictor_and_reorder(buf);
        } else if (compression == 2 || compression == 3) {
            buf.resize(expect);
            uLongf dl = (uLongf)expect;
            if (uncompress(buf.data(), &dl, src, (uLong)size) != Z_OK || dl != expect) {
                err = "parse: bad EXR ZIP chunk";
                return false;
            }
            exr_undo_predictor_and_reorder(buf);
        } else {
            err = "parse: EXR chunk size mismatch";
            return false;
        }
        size_t q = 0;
        for (int64_t yy = 0; yy < ny; ++yy) {
            float *row = &rgba[(size_t)(ys + yy) * W * 4];
            for (size_t c = 0; c < chans.size(); ++c) {
                int slot = -1;
                for (int s = 0; s < 4; ++s)
                    if (ci[s] == (int)c) slot = s;
                const bool gray = (int)c == cy && (ci[0] < 0 || ci[1] < 0 || ci[2] < 0);
                for (int64_t x = 0; x < W; ++x) {
                    float v;
                    if (chans[c].type == 1) {
                        uint16_t hv;
                        std::memcpy(&hv, &buf[q], 2);
                        v = half_to_float(hv);
                        q += 2;
                    } else if (chans[c].type == 2) {
                        std::memcpy(&v, &buf[q], 4);
                        q += 4;
                    } else {
                        uint32_t u;
                        std::memcpy(&u, &buf[q], 4);
                        v = (float)u;
                        q += 4;
                    }
                    if (slot >= 0) row[4 * x + slot] = v;
                    if (gray) row[4 * x] = row[4 * x + 1] = row[4 * x + 2] = v;
                }
            }
        }
    }
    return true;
}

// ------------------------------------------------------------------ Radiance RGBE
bool load_hdr(const std::vector<unsigned char> &f, uint32_t &w, uint32_t &h, std::vector<float> &rgba,
              std::string &err) {
    size_t p = 0;
    auto line = [&](std::string &s) {
        s.clear();
        while (p < f.size() && f[p] != '\n') s.push_back((char)f[p++]);
        if (p >= f.size()) return false;
        ++p;
        return true;
    };
    std::string s;
    if (!line(s) || (s.rfind("#?RADIANCE", 0) != 0 && s.rfind("#?RGBE", 0) != 0)) {
        err = "parse: not a Radiance HDR file";
        return false;
    }
    bool fmt_ok = true;
    for (;;) {
        if (!line(s)) {
            err = "parse: truncated HDR header";
            return false;
        }
        if (s.empty()) break;
        if (s.rfind("FORMAT=", 0) == 0 && s != "FORMAT=32-bit_rle_rgbe") fmt_ok = false;
    }
    if (!fmt_ok) {
        err = "unsupported: HDR pixel format other than 32-bit_rle_rgbe";
        return false;
    }
    if (!line(s)) {
        err = "parse: missing HDR resolution";
        return false;
    }
    char ya[3] = {0}, xa[3] = {0};
    int H = 0, W = 0;
    if (std::sscanf(s.c_str(), "%2s %d %2s %d", ya, &H, xa, &W) != 4 || std::string(xa) != "+X" ||
        (std::string(ya) != "-Y" && std::string(ya) != "+Y") || W <= 0 || H <= 0 || W > 65536 || H > 65536) {
        err = "unsupported: HDR orientation other than -Y H +X W / +Y H +X W";
        return false;
    }
    if (!image_size_ok((uint64_t)W, (uint64_t)H, err)) return false;
    const bool bottom_up = std::string(ya) == "+Y";
    w = (uint32_t)W;
    h = (uint32_t)H;
    rgba.assign((size_t)W * H * 4, 0.f);
    std::vector<unsigned char> scan((size_t)W * 4);
    for (int y = 0; y < H; ++y) {
        if (p + 4 > f.size()) {
            err = "parse: truncated HDR data";
            return false;
        }
        const bool rle = W >= 8 && W < 32768 && f[p] == 2 && f[p + 1] == 2 && ((f[p + 2] << 8) | f[p + 3]) == W &&
                         !(f[p + 2] & 0x80);
        if (rle) {
            p += 4;
            for (int c = 0; c < 4; ++c) {
                int x = 0;
                while (x < W) {
                    if (p >= f.size()) {
                        err = "parse: truncated HDR run";
                        return false;
                    }
                    int n = f[p++];
                    if (n > 128) {
                        n -= 128;
                        if (x + n > W || p >= f.size()) {
                            err = "parse: bad HDR run";
                            return false;
                        }
                        const unsigned char v = f[p++];
                        for (int k = 0; k < n; ++k) scan[4 * (size_t)(x++) + c] = v;
                    } else {
                        if (n == 0 || x + n > W || p + (size_t)n > f.size()) {
                            err = "parse: bad HDR run";
                            return false;
                        }
                        for (int k = 0; k < n; ++k) scan[4 * (size_t)(x++) + c] = f[p++];
                    }
                }
            }
        } else {
            if (p + (size_t)W * 4 > f.size()) {
                err = "parse: truncated HDR data";
                return false;
            }
            std::memcpy(scan.data(), &f[p], (size_t)W * 4);
            p += (size_t)W * 4;
        }
        float *row = &rgba[(size_t)(bottom_up ? H - 1 - y : y) * W * 4];
        for (int x = 0; x < W; ++x) {
            const unsigned char *e = &scan[4 * (size_t)x];
            const float m = e[3] ? std::ldexp(1.f, (int)e[3] - 136) : 0.f;
            row[4 * x + 0] = e[0] * m;
            row[4 * x + 1] = e[1] * m;
            row[4 * x + 2] = e[2] * m;
            row[4 * x + 3] = 1.f;
        }
    }
    return true;
}

// ------------------------------------------------------------------ PNG
uint32_t be32(const unsigned char *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

bool load_png(const std::vector<unsigned char> &f, uint32_t &w, uint32_t &h, std::vector<float> &rgba,
              std::string &err) {
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) {
        err = "parse: not a PNG file";
        return false;
    }
    size_t p = 8;
    uint32_t W = 0, H = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<unsigned char> idat, plte;
    bool have_ihdr = false;
    while (p + 8 <= f.size()) {
        const uint32_t len = be32(&f[p]);
        const std::string type(reinterpret_cast<const char *>(&f[p + 4]), 4);
        p += 8;
        if (p + (size_t)len + 4 > f.size()) {
            err = "parse: truncated PNG chunk";
            return false;
        }
        const unsigned char *d = &f[p];
        if (type == "IHDR" && len >= 13) {
            W = be32(d);
            H = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
            have_ihdr = true;
        } else if (type == "PLTE") {
            plte.assign(d, d + len);
        } else if (type == "IDAT") {
            idat.insert(idat.end(), d, d + len);
        } else if (type == "IEND") {
            break;
        }
        p += (size_t)len + 4;  // data + CRC
    }
    if (!have_ihdr || W == 0 || H == 0 || W > 65536 || H > 65536) {
        err = "parse: PNG without a valid IHDR";
        return false;
    }
    if (!image_size_ok(W, H, err)) return false;
    if (interlace > 1) {
        err = "parse: bad PNG interlace method";
        return false;
    }
    int chans;
    switch (ctype) {
        case 0: chans = 1; break;
        case 2: chans = 3; break;
        case 3: chans = 1; break;
        case 4: chans = 2; break;
        case 6: chans = 4; break;
        default:
            err = "parse: bad PNG color type";
            return false;
    }
    const bool depth_ok = depth == 8 || (depth == 16 && ctype != 3) || ((ctype == 0 || ctype == 3) && depth < 8 &&
                                                                         (depth == 1 || depth == 2 || depth == 4));
    if (!depth_ok || (ctype == 3 && plte.size() < 3)) {
        err = "parse: bad PNG bit depth or palette";
        return false;
    }
    // Passes: one for a plain image; Adam7 (PNG spec 8.2) splits the image
    // into seven sub-images (x0, y0, dx, dy), each filtered on its own, stored
    // one after the other; an empty pass has no bytes at all.
    struct Pass { uint32_t x0, y0, dx, dy, w, h; };
    static const uint32_t adam7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                         {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    std::vector<Pass> passes;
    for (int k = 0; k < (interlace ? 7 : 1); ++k) {
        Pass ps{0, 0, 1, 1, W, H};
        if (interlace) {
            ps = Pass{adam7[k][0], adam7[k][1], adam7[k][2], adam7[k][3], 0, 0};
            ps.w = W > ps.x0 ? (W - ps.x0 + ps.dx - 1) / ps.dx : 0;
            ps.h = H > ps.y0 ? (H - ps.y0 + ps.dy - 1) / ps.dy : 0;
        }
        if (ps.w && ps.h) passes.push_back(ps);
    }
    const size_t bits_pp = (size_t)chans * depth;
    const size_t bpp = std::max<size_t>(1, bits_pp / 8);
    auto row_bytes_of = [&](uint32_t pw) { return ((size_t)pw * bits_pp + 7) / 8; };
    size_t total = 0;
    for (const Pass &ps : passes) total += (row_bytes_of(ps.w) + 1) * ps.h;
    std::vector<unsigned char> raw(total);
    uLongf dl = (uLongf)raw.size();
    if (idat.empty() || uncompress(raw.data(), &dl, idat.data(), (uLong)idat.size()) != Z_OK || dl != raw.size()) {
        err = "parse: bad PNG image data";
        return false;
    }
    w = W;
    h = H;
    rgba.assign((size_t)W * H * 4, 0.f);
    float lut[256];
    for (int i = 0; i < 256; ++i) lut[i] = std::pow(i * 1.f / 255.f, 2.2f);
    size_t off = 0;
    for (const Pass &ps : passes) {
        const size_t row_bytes = row_bytes_of(ps.w);
        std::vector<unsigned char> prev(row_bytes, 0), cur(row_bytes);
        for (uint32_t y = 0; y < ps.h; ++y) {
            const unsigned char *src = &raw[off + (row_bytes + 1) * y];
            const int ft = src[0];
            for (size_t i = 0; i < row_bytes; ++i) {
                const int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
                int v = src[1 + i];
                switch (ft) {
                    case 0: break;
                    case 1: v += a; break;
                    case 2: v += b; break;
                    case 3: v += (a + b) / 2; break;
                    case 4: {
                        const int pp = a + b - c, pa = std::abs(pp - a), pb = std::abs(pp - b), pc = std::abs(pp - c);
                        v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                    } break;
                    default:
                        err = "parse: bad PNG filter";
                        return false;
                }
                cur[i] = (unsigned char)v;
            }
            float *row = &rgba[(size_t)(ps.y0 + y * ps.dy) * W * 4];
            for (uint32_t x = 0; x < ps.w; ++x) {
                unsigned char px[4] = {0, 0, 0, 255};
                if (depth < 8) {
                    const size_t bit = (size_t)x * depth;
                    const int v = (cur[bit / 8] >> (8 - depth - (int)(bit % 8))) & ((1 << depth) - 1);
                    if (ctype == 3) {
                        if ((size_t)v * 3 + 2 >= plte.size()) {
                            err = "parse: PNG palette index out of range";
                            return false;
                        }
                        for (int k = 0; k < 3; ++k) px[k] = plte[3 * v + k];
                    } else {
                        const int scale = depth == 1 ? 255 : (depth == 2 ? 0x55 : 0x11);
                        px[0] = px[1] = px[2] = (unsigned char)(v * scale);
                    }
                } else {
                    const size_t step = depth / 8;
                    unsigned char sm[4];
                    for (int k = 0; k < chans; ++k) sm[k] = cur[((size_t)x * chans + k) * step];  // high byte
                    if (ctype == 3) {
                        if ((size_t)sm[0] * 3 + 2 >= plte.size()) {
                            err = "parse: PNG palette index out of range";
                            return false;
                        }
                        for (int k = 0; k < 3; ++k) px[k] = plte[3 * (size_t)sm[0] + k];
                    } else if (chans <= 2) {
                        px[0] = px[1] = px[2] = sm[0];
                        if (chans == 2) px[3] = sm[1];
                    } else {
                        for (int k = 0; k < chans; ++k) px[k] = sm[k];
                    }
                }
                float *o = row + 4 * (size_t)(ps.x0 + x * ps.dx);
                o[0] = lut[px[0]];
                o[1] = lut[px[1]];
                o[2] = lut[px[2]];
                o[3] = px[3] * 1.f / 255.f;
            }
            prev.swap(cur);
        }
        off += (row_bytes + 1) * ps.h;
    }
    return true;
}

// 8-bit pixels (1-4 channels, row 0 = top) to linear RGBA float: color
// channels (v / 255)^2.2, alpha v / 255; gray is expanded to RGB.
void rgba_from_8bit(const std::vector<unsigned char> &px, int chans, uint32_t W, uint32_t H, std::vector<float> &rgba) {
    float lut[256];
    for (int i = 0; i < 256; ++i) lut[i] = std::pow(i * 1.f / 255.f, 2.2f);
    rgba.assign((size_t)W * H * 4, 0.f);
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        const unsigned char *s = &px[i * chans];
        const unsigned char r = s[0], g = chans >= 3 ? s[1] : s[0], b = chans >= 3 ? s[2] : s[0];
        const unsigned char a = chans == 4 ? s[3] : (chans == 2 ? s[1] : 255);
        rgba[4 * i + 0] = lut[r];
        rgba[4 * i + 1] = lut[g];
        rgba[4 * i + 2] = lut[b];
        rgba[4 * i + 3] = a * 1.f / 255.f;
    }
}

uint32_t le16(const unsigned char *p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8; }
uint32_t le32(const unsigned char *p) { return le16(p) | le16(p + 2) << 16; }

// ------------------------------------------------------------------ PSD
// The composite (merged) image of a Photoshop file, as stb_image reads it:
// RGB color mode, 8 or 16 bits per channel (16-bit samples keep their high
// byte), raw or PackBits (RLE) compressed, channel-planar; a fourth channel
// is alpha, a file with fewer than four channels is opaque.  Layers, masks
// and image resources are skipped.
bool load_psd(const std::vector<unsigned char> &f, uint32_t &w, uint32_t &h, std::vector<unsigned char> &px,
              std::string &err) {
    auto u16 = [&](size_t at) { return (uint32_t)f[at] << 8 | f[at + 1]; };
    auto u32 = [&](size_t at) { return u16(at) << 16 | u16(at + 2); };
    if (f.size() < 26 || std::memcmp(f.data(), "8BPS", 4) != 0 || u16(4) != 1) {
        err = "parse: not a PSD file";
        return false;
    }
    const uint32_t nch = u16(12), H = u32(14), W = u32(18), depth = u16(22), mode = u16(24);
    if (nch < 1 || nch > 56 || W == 0 || H == 0 || W > 30000 || H > 30000) {
        err = "parse: bad PSD header";
        return false;
    }
    if (mode != 3 || (depth != 8 && depth != 16)) {
        err = "unsupported: PSD color mode or depth (8/16-bit RGB is read)";
        return false;
    }
    if (!image_size_ok(W, H, err)) return false;
    size_t p = 26;
    for (int section = 0; section < 3; ++section) {  // color mode data, image resources, layers and masks
        if (p + 4 > f.size()) {
            err = "parse: truncated PSD file";
            return false;
        }
        p += 4 + (size_t)u32(p);
    }
    if (p + 2 > f.size()) {
        err = "parse: truncated PSD file";
        return false;
    }
    const uint32_t comp = u16(p);
    p += 2;
    if (comp > 1) {
        err = "unsupported: PSD compression (raw and PackBits are read)";
        return false;
    }
    const size_t bps = depth / 8, row = (size_t)W * bps, plane = row * H;
    const uint32_t used = std::min<uint32_t>(nch, 4);
    // Every row costs at least one stored byte (raw: all of it), so a file
    // shorter than that is truncated whatever its header claims.
    if (p + (comp == 0 ? (size_t)used * plane : (size_t)2 * nch * H + (size_t)used * H) > f.size()) {
        err = "parse: truncated PSD image data";
        return false;
    }
    if (comp == 1) {
        // PackBits stores at most 128 bytes per two-byte packet, so a row of
        // `row` bytes needs at least 2 ceil(row / 128) stored bytes; the
        // declared counts of the rows read must all be that large and fit in
        // the file before the planes are allocated.
        const size_t min_row = 2 * ((row + 127) / 128);
        size_t total = 0;
        for (uint32_t c = 0; c < used; ++c)
            for (uint32_t y = 0; y < H; ++y) {
                const size_t n = u16(p + 2 * ((size_t)c * H + y));
                if (n < min_row) {
                    err = "parse: PSD PackBits row shorter than its width allows";
                    return false;
                }
                total += n;
            }
        if (p + (size_t)2 * nch * H + total > f.size()) {
            err = "parse: truncated PSD image data";
            return false;
        }
    }
    std::vector<unsigned char> planes((size_t)used * plane);
    if (comp == 0) {
        std::memcpy(planes.data(), &f[p], planes.size());
    } else {
        // Byte counts of every row of every channel, then the PackBits rows.
        size_t q = p + (size_t)2 * nch * H;
        if (q > f.size()) {
            err = "parse: truncated PSD image data";
            return false;
        }
        for (uint32_t c = 0; c < used; ++c)
            for (uint32_t y = 0; y < H; ++y) {
                const size_t n = u16(p + 2 * ((size_t)c * H + y)), stop = q + n;
                if (stop > f.size()) {
                    err = "parse: truncated PSD image data";
                    return false;
                }
                unsigned char *dst = &planes[(size_t)c * plane + (size_t)y * row];
                size_t o = 0;
                while (q < stop && o < row) {
                    const int hdr = (signed char)f[q++];
                    if (hdr >= 0) {  // hdr + 1 literal bytes
                        const size_t m = std::min<size_t>((size_t)hdr + 1, std::min(row - o, stop - q));
                        std::memcpy(dst + o, &f[q], m);
                        o += m;
                        q += (size_t)hdr + 1;
                    } else if (hdr != -128) {  // one byte repeated 1 - hdr times
                        if (q >= stop) break;
                        const size_t m = std::min<size_t>((size_t)(1 - hdr), row - o);
                        std::memset(dst + o, f[q++], m);
                        o += m;
                    }
                }
                if (o != row) {
                    err = "parse: bad PSD PackBits row";
                    return false;
                }
                q = stop;
            }
    }
    w = W;
    h = H;
    px.assign((size_t)W * H * 4, 255);
    for (uint32_t c = 0; c < used; ++c)
        for (size_t i = 0; i < (size_t)W * H; ++i) px[4 * i + c] = planes[(size_t)c * plane + i * bps];  // high byte
    if (used < 3)
        for (size_t i = 0; i < (size_t)W * H; ++i) px[4 * i + 1] = px[4 * i + 2] = px[4 * i];
    return true;
}

// ------------------------------------------------------------------ GIF
// First image of a GIF87a / GIF89a file, as stb_image returns it: a canvas of
// the logical screen size, transparent black where the image does not cover
// it or where the graphic control extension marks the transparent index;
// colors from the local table, else the global one.  LZW (variable code
// width from min code size + 1 up to 12 bits, clear and end codes, the table
// kept full once 4096 codes exist) and the four-pass row order of an
// interlaced image follow the GIF89a specification.
bool load_gif(const std::vector<unsigned char> &f, uint32_t &w, uint32_t &h, std::vector<unsigned char> &px,
              std::string &err) {
    if (f.size() < 13 || (std::memcmp(f.data(), "GIF87a", 6) != 0 && std::memcmp(f.data(), "GIF89a", 6) != 0)) {
        err = "parse: not a GIF file";
        return false;
    }
    const uint32_t W = f[6] | f[7] << 8, H = f[8] | f[9] << 8;
    if (W == 0 || H == 0) {
        err = "parse: GIF with an empty screen";
        return false;
    }
    if (!image_size_ok(W, H, err)) return false;
    size_t p = 13;
    std::vector<unsigned char> gct;
    if (f[10] & 0x80) {
        const size_t n = (size_t)3 << ((f[10] & 7) + 1);
        if (p + n > f.size()) {
            err = "parse: truncated GIF color table";
            return false;
        }
        gct.assign(f.begin() + (std::ptrdiff_t)p, f.begin() + (std::ptrdiff_t)(p + n));
        p += n;
    }
    int transparent = -1;
    auto skip_sub_blocks = [&]() -> bool {
        while (p < f.size()) {
            const size_t n = f[p++];
            if (n == 0) return true;
            p += n;
        }
        return false;
    };
    while (p < f.size()) {
        const unsigned char b = f[p++];
        if (b == 0x3B) break;  // trailer
        if (b == 0x21) {       // extension
            if (p >= f.size()) break;
            const unsigned char label = f[p++];
            if (label == 0xF9 && p + 5 <= f.size() && f[p] >= 4) {  // graphic control
                transparent = (f[p + 1] & 1) ? f[p + 4] : -1;
            }
            if (!skip_sub_blocks()) break;
            continue;
        }
        if (b != 0x2C) {
            err = "parse: bad GIF block";
            return false;
        }
        if (p + 9 > f.size()) break;
        const uint32_t ix = f[p] | f[p + 1] << 8, iy = f[p + 2] | f[p + 3] << 8;
        const uint32_t iw = f[p + 4] | f[p + 5] << 8, ih = f[p + 6] | f[p + 7] << 8;
        const unsigned char flags = f[p + 8];
        p += 9;
        if ((uint64_t)iw * ih > (uint64_t)W * H || iw == 0 || ih == 0) {
            err = "parse: GIF image larger than its screen";
            return false;
        }
        std::vector<unsigned char> lct;
        if (flags & 0x80) {
            const size_t n = (size_t)3 << ((flags & 7) + 1);
            if (p + n > f.size()) break;
            lct.assign(f.begin() + (std::ptrdiff_t)p, f.begin() + (std::ptrdiff_t)(p + n));
            p += n;
        }
        const std::vector<unsigned char> &ct = lct.empty() ? gct : lct;
        if (ct.empty() || p >= f.size()) {
            err = "parse: GIF image without a color table";
            return false;
        }
        const int min_size = f[p++];
        if (min_size < 2 || min_size > 11) {
            err = "parse: bad GIF LZW code size";
            return false;
        }
        // Data sub-blocks joined into one stream.
        std::vector<unsigned char> data;
        while (p < f.size()) {
            const size_t n = f[p++];
            if (n == 0) break;
            if (p + n > f.size()) {
                err = "parse: truncated GIF image data";
                return false;
            }
            data.insert(data.end(), f.begin() + (std::ptrdiff_t)p, f.begin() + (std::ptrdiff_t)(p + n));
            p += n;
        }
        // LZW: code k < clear is the literal k; each new code is an earlier
        // code's string plus one byte.
        const int clear = 1 << min_size, end = clear + 1;
        std::vector<uint16_t> prefix(4096);
        std::vector<unsigned char> suffix(4096), first(4096), stack;
        for (int k = 0; k < clear; ++k) {
            suffix[(size_t)k] = (unsigned char)k;
            first[(size_t)k] = (unsigned char)k;
        }
        std::vector<unsigned char> idx;
        idx.reserve((size_t)iw * ih);
        int width = min_size + 1, next = end + 1, prev = -1;
        uint32_t acc = 0;
        int nacc = 0;
        size_t q = 0;
        while (idx.size() < (size_t)iw * ih) {
            while (nacc < width && q < data.size()) {
                acc |= (uint32_t)data[q++] << nacc;
                nacc += 8;
            }
            if (nacc < width) break;  // data ended early: the rest stays uncovered
            const int code = (int)(acc & ((1u << width) - 1u));
            acc >>= width;
            nacc -= width;
            if (code == clear) {
                width = min_size + 1;
                next = end + 1;
                prev = -1;
                continue;
            }
            if (code == end) break;
            int c = code;
            if (code > next || (code == next && prev < 0)) {
                err = "parse: bad GIF LZW code";
                return false;
            }
            stack.clear();
            if (code == next) {  // the string being defined: prev's string + its first byte
                stack.push_back(first[(size_t)prev]);
                c = prev;
            }
            while (c >= clear) {
                stack.push_back(suffix[(size_t)c]);
                c = prefix[(size_t)c];
            }
            stack.push_back((unsigned char)c);
            for (auto it = stack.rbegin(); it != stack.rend() && idx.size() < (size_t)iw * ih; ++it) idx.push_back(*it);
            if (prev >= 0 && next < 4096) {
                prefix[(size_t)next] = (uint16_t)prev;
                suffix[(size_t)next] = (unsigned char)c;  // first byte of this code's string
                first[(size_t)next] = first[(size_t)prev];
                ++next;
                if (next == (1 << width) && width < 12) ++width;
            }
            prev = code;
        }
        // Rows in file order: top to bottom, or the four interlace passes.
        std::vector<uint32_t> rows;
        if (flags & 0x40) {
            static const uint32_t start[4] = {0, 4, 2, 1}, step[4] = {8, 8, 4, 2};
            for (int ps = 0; ps < 4; ++ps)
                for (uint32_t y = start[ps]; y < ih; y += step[ps]) rows.push_back(y);
        } else {
            for (uint32_t y = 0; y < ih; ++y) rows.push_back(y);
        }
        w = W;
        h = H;
        px.assign((size_t)W * H * 4, 0);
        for (size_t r = 0; r < rows.size(); ++r) {
            const uint32_t y = iy + rows[r];
            for (uint32_t x = 0; x < iw; ++x) {
                const size_t k = r * iw + x;
                if (k >= idx.size()) break;
                const int v = idx[k];
                if (v == transparent || y >= H || ix + x >= W || (size_t)v * 3 + 2 >= ct.size()) continue;
                unsigned char *o = &px[((size_t)y * W + ix + x) * 4];
                o[0] = ct[(size_t)v * 3];
                o[1] = ct[(size_t)v * 3 + 1];
                o[2] = ct[(size_t)v * 3 + 2];
                o[3] = 255;
            }
        }
        return true;
    }
    err = "parse: GIF without an image";
    return false;
}

// ------------------------------------------------------------------ BMP
// Windows bitmap: BITMAPINFOHEADER and later headers, uncompressed 1/4/8-bit
// palette and 24/32-bit pixels, 16/32-bit BI_BITFIELDS; bottom-up or top-down.
bool load_bmp(const std::vector<unsigned char> &f, uint32_t &w, uint32_t &h, std::vector<float> &rgba,
              std::string &err) {
    if (f.size() < 54 || f[0] != 'B' || f[1] != 'M') {
        err = "parse: not a BMP file";
        return false;
    }
    const uint32_t off = le32(&f[10]), hsize = le32(&f[14]);
    if (hsize < 40 || 14 + (size_t)hsize > f.size()) {
        err = "unsupported: BMP header";
        return false;
    }
    const int32_t W = (int32_t)le32(&f[18]), Hs = (int32_t)le32(&f[22]);
    const uint32_t bpp = le16(&f[28]), comp = le32(&f[30]);
    uint32_t ncolors = le32(&f[46]);
    const bool top_down = Hs < 0;
    const uint32_t H = (uint32_t)(top_down ? -(int64_t)Hs : Hs);
    if (W <= 0 || H == 0 || W > 65536 || H > 65536) {
        err = "parse: bad BMP size";
        return false;
    }
    uint32_t mask[4] = {0x00FF0000u, 0x0000FF00u, 0x000000FFu, 0u};
    if (bpp == 16) {
        mask[0] = 0x7C00u;
        mask[1] = 0x03E0u;
        mask[2] = 0x001Fu;
    }
    if (comp == 3 || comp == 6) {  // BI_BITFIELDS / BI_ALPHABITFIELDS
        const size_t mp = hsize >= 52 ? 54 : 14 + hsize;
        if (bpp != 16 && bpp != 32) {
            err = "unsupported: BMP bit fields at this depth";
            return false;
        }
        if (mp + 12 > f.size()) {
            err = "parse: truncated BMP";
            return false;
        }
        for (int k = 0; k < 3; ++k) mask[k] = le32(&f[mp + 4 * k]);
        mask[3] = (hsize >= 56 || comp == 6) && mp + 16 <= f.size() ? le32(&f[mp + 12]) : 0u;
    } else if (comp != 0) {
        err = "unsupported: compressed BMP";
        return false;
    } else if (bpp == 32) {
        mask[3] = 0u;  // plain 32-bit BMP: the fourth byte is not alpha
    }
    if (bpp != 1 && bpp != 4 && bpp != 8 && bpp != 16 && bpp != 24 && bpp != 32) {
        err = "unsupported: BMP bit depth";
        return false;
    }
    std::vector<unsigned char> pal;
    if (bpp <= 8) {
        if (ncolors == 0) ncolors = 1u << bpp;
        const size_t pp = 14 + hsize;
        if (ncolors > 256 || pp + 4 * (size_t)ncolors > f.size()) {
            err = "parse: bad BMP palette";
            return false;
        }
        pal.assign(&f[pp], &f[pp] + 4 * (size_t)ncolors);
    }
    const size_t stride = (((size_t)W * bpp + 31) / 32) * 4;
    if ((size_t)off + stride * H > f.size()) {
        err = "parse: truncated BMP";
        return false;
    }
    auto field = [](uint32_t v, uint32_t m) -> unsigned char {
        if (m == 0) return 255;
        int shift = 0;
        while (!((m >> shift) & 1u)) ++shift;
        const uint32_t bits = m >> shift;
        const uint32_t x = (v & m) >> shift;
        return (unsigned char)((x * 255u + bits / 2) / bits);
    };
    const bool has_alpha = mask[3] != 0;
    const int chans = has_alpha ? 4 : 3;
    if (!image_size_ok((uint64_t)W, (uint64_t)H, err)) return false;
    std::vector<unsigned char> px((size_t)W * H * chans);
    for (uint32_t y = 0; y < H; ++y) {
        const unsigned char *row = &f[off + stride * (top_down ? y : H - 1 - y)];
        for (int32_t x = 0; x < W; ++x) {
            unsigned char *d = &px[((size_t)y * W + x) * chans];
            if (bpp <= 8) {
                const size_t bit = (size_t)x * bpp;
                const uint32_t idx = (row[bit / 8] >> (8 - bpp - (bit % 8))) & ((1u << bpp) - 1u);
                if (idx >= ncolors) {
                    err = "parse: BMP palette index out of range";
                    return false;
                }
                d[0] = pal[4 * idx + 2];
                d[1] = pal[4 * idx + 1];
                d[2] = pal[4 * idx + 0];
            } else if (bpp == 24) {
                d[0] = row[3 * x + 2];
                d[1] = row[3 * x + 1];
                d[2] = row[3 * x + 0];
            } else {
                const uint32_t v = bpp == 16 ? le16(&row[2 * x]) : le32(&row[4 * x]);
                for (int k = 0; k < 3; ++k) d[k] = field(v, mask[k]);
                if (has_alpha) d[3] = field(v, mask[3]);
            }
        }
    }
    w = (uint32_t)W;
    h = H;
    rgba_from_8bit(px, chans, w, h, rgba);
    return true;
}

// ------------------------------------------------------------------ TGA
// Truevision TGA: color-mapped (1, 9), true-color (2, 10) and gray (3, 11)
// images, raw or run-length encoded, 8/15/16/24/32 bits; origin bit honored.
bool load_tga(const std::vector<unsigned char> &f, uint32_t &w, uint32_t &h, std::vector<float> &rgba,
              std::string &err) {
    if (f.size() < 18) {
        err = "parse: not a TGA file";
        return false;
    }
    const uint32_t id_len = f[0], cmap_type = f[1], type = f[2];
    const uint32_t cmap_first = le16(&f[3]), cmap_len = le16(&f[5]), cmap_bits = f[7];
    const uint32_t W = le16(&f[12]), H = le16(&f[14]), bpp = f[16], desc = f[17];
    const bool rle = type >= 9;
    const uint32_t base = rle ? type - 8 : type;
    if ((base != 1 && base != 2 && base != 3) || W == 0 || H == 0) {
        err = "unsupported: TGA image type";
        return false;
    }
    if ((base == 1 && (bpp != 8 || cmap_type != 1)) || (base == 2 && bpp != 15 && bpp != 16 && bpp != 24 && bpp != 32) ||
        (base == 3 && bpp != 8)) {
        err = "unsupported: TGA pixel depth";
        return false;
    }
    size_t p = 18 + id_len;
    const uint32_t cbytes = (cmap_bits + 7) / 8;
    std::vector<unsigned char> cmap;
    if (cmap_type == 1) {
        if (p + (size_t)cmap_len * cbytes > f.size()) {
            err = "parse: truncated TGA color map";
            return false;
        }
        cmap.assign(&f[p], &f[p] + (size_t)cmap_len * cbytes);
        p += (size_t)cmap_len * cbytes;
    }
    const bool alpha = (base == 1 ? cmap_bits : bpp) == 32;
    const int chans = base == 3 ? 1 : (alpha ? 4 : 3);
    auto decode = [&](const unsigned char *s, unsigned char *d) -> bool {
        const unsigned char *c = s;
        uint32_t nb = (bpp + 7) / 8;
        if (base == 1) {
            const uint32_t idx = s[0];
            if (idx < cmap_first || idx - cmap_first >= cmap_len) return false;
            c = &cmap[(size_t)(idx - cmap_first) * cbytes];
            nb = cbytes;
        }
        if (base == 3) {
            d[0] = c[0];
        } else if (nb == 2) {
            const uint32_t v = le16(c);
            d[0] = (unsigned char)((((v >> 10) & 31u) * 255u + 15u) / 31u);
            d[1] = (unsigned char)((((v >> 5) & 31u) * 255u + 15u) / 31u);
            d[2] = (unsigned char)(((v & 31u) * 255u + 15u) / 31u);
        } else {
            d[0] = c[2];
            d[1] = c[1];
            d[2] = c[0];
            if (chans == 4) d[3] = c[3];
        }
        return true;
    };
    const uint32_t sbytes = (bpp + 7) / 8;  // bytes per stored pixel
    if (!image_size_ok((uint64_t)W, (uint64_t)H, err)) return false;
    std::vector<unsigned char> px((size_t)W * H * chans);
    const size_t n = (size_t)W * H;
    size_t i = 0;
    while (i < n) {
        uint32_t count = 1;
        bool run = false;
        if (rle) {
            if (p >= f.size()) break;
            const unsigned char hdr = f[p++];
            count = (hdr & 0x7Fu) + 1u;
            run = (hdr & 0x80u) != 0;
        }
        for (uint32_t k = 0; k < count && i < n; ++k, ++i) {
            const size_t src = run ? p : p + (size_t)k * sbytes;
            if (src + sbytes > f.size() || !decode(&f[src], &px[i * chans])) {
                err = "parse: truncated or bad TGA pixel data";
                return false;
            }
        }
        p += run ? sbytes : (size_t)count * sbytes;
        if (!rle) {
            p = 18 + id_len + cmap.size() + i * sbytes;
        }
    }
    if (i < n) {
        err = "parse: truncated TGA pixel data";
        return false;
    }
    // Origin: bit 5 set = top-left, else bottom-left; bit 4 = right-to-left.
    std::vector<unsigned char> out(px.size());
    const bool top = (desc & 0x20u) != 0, right = (desc & 0x10u) != 0;
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            const uint32_t sy = top ? y : H - 1 - y, sx = right ? W - 1 - x : x;
            std::memcpy(&out[((size_t)y * W + x) * chans], &px[((size_t)sy * W + sx) * chans], chans);
        }
    w = W;
    h = H;
    rgba_from_8bit(out, chans, w, h, rgba);
    return true;
}

// ------------------------------------------------------------------ PNM
// Binary netpbm: P5 (gray) and P6 (RGB), maxval up to 65535 (high byte kept).
bool load_pnm(const std::vector<unsigned char> &f, uint32_t &w, uint32_t &h, std::vector<float> &rgba,
              std::string &err) {
    if (f.size() < 3 || f[0] != 'P' || (f[1] != '5' && f[1] != '6')) {
        err = "unsupported: PNM variant (P5 and P6 are read)";
        return false;
    }
    const int chans = f[1] == '6' ? 3 : 1;
    size_t p = 2;
    uint32_t v[3];
    for (int k = 0; k < 3; ++k) {
        while (p < f.size()) {
            if (f[p] == '#') {
                while (p < f.size() && f[p] != '\n') ++p;
            } else if (std::isspace(f[p])) {
                ++p;
            } else {
                break;
            }
        }
        uint64_t x = 0;
        size_t digits = 0;
        while (p < f.size() && std::isdigit(f[p]) && digits < 10) x = x * 10 + (f[p++] - '0'), ++digits;
        if (digits == 0 || x == 0 || x > 65535) {
            err = "parse: bad PNM header";
            return false;
        }
        v[k] = (uint32_t)x;
    }
    ++p;  // the single whitespace byte before the raster
    const uint32_t W = v[0], H = v[1], maxv = v[2];
    const size_t bpc = maxv > 255 ? 2 : 1;
    if (p + (size_t)W * H * chans * bpc > f.size()) {
        err = "parse: truncated PNM raster";
        return false;
    }
    if (!image_size_ok((uint64_t)W, (uint64_t)H, err)) return false;
    std::vector<unsigned char> px((size_t)W * H * chans);
    for (size_t i = 0; i < px.size(); ++i) {
        const uint32_t s = bpc == 2 ? ((uint32_t)f[p + 2 * i] << 8 | f[p + 2 * i + 1]) : f[p + i];
        px[i] = (unsigned char)(((uint64_t)s * 255u + maxv / 2) / maxv);
    }
    w = W;
    h = H;
    rgba_from_8bit(px, chans, w, h, rgba);
    return true;
}

}  // namespace

bool load_jpeg(const std::vector<unsigned char> &f, uint32_t &w, uint32_t &h, std::vector<unsigned char> &px,
               int &chans, std::string &err);

bool load_image(const std::string &path, uint32_t &w, uint32_t &h, std::vector<float> &rgba, std::string &err) {
    std::vector<unsigned char> f;
    if (!read_file(path, f, err)) return false;
    const std::string e = lower_ext(path);
    bool ok;
    if (e == "exr") ok = load_exr(f, w, h, rgba, err);
    else if (e == "hdr" || e == "pic") ok = load_hdr(f, w, h, rgba, err);
    else if (e == "png") ok = load_png(f, w, h, rgba, err);
    else if (e == "bmp") ok = load_bmp(f, w, h, rgba, err);
    else if (e == "tga") ok = load_tga(f, w, h, rgba, err);
    else if (e == "ppm" || e == "pgm" || e == "pnm") ok = load_pnm(f, w, h, rgba, err);
    else if (e == "psd") {
        std::vector<unsigned char> px;
        ok = load_psd(f, w, h, px, err);
        if (ok) rgba_from_8bit(px, 4, w, h, rgba);
    } else if (e == "gif") {
        std::vector<unsigned char> px;
        ok = load_gif(f, w, h, px, err);
        if (ok) rgba_from_8bit(px, 4, w, h, rgba);
    } else if (e == "jpg" || e == "jpeg") {
        std::vector<unsigned char> px;
        int chans = 0;
        ok = load_jpeg(f, w, h, px, chans, err);
        if (ok) rgba_from_8bit(px, chans, w, h, rgba);
    } else {
        err = "unsupported: image format ." + e + " (EXR, HDR, PNG, JPEG, GIF, PSD, BMP, TGA and PNM are read)";
        return false;
    }
    if (!ok && err.find(path) == std::string::npos) err += " (" + path + ")";
    return ok;
}

}  // namespace pupil
