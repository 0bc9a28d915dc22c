// JPEG reader for bitmap textures (util/texture.cpp:87-116 reads JPEG files
// through stb_image).  Written from ITU-T T.81 / JFIF 1.02: sequential
// (SOF0 / SOF1) and progressive (SOF2, Annex G: spectral selection and
// successive approximation) DCT with Huffman coding, 8-bit samples, 1 (gray)
// or 3 (YCbCr) components with sampling factors up to 4, restart intervals,
// scans holding any subset of the components.  Lossless and arithmetic-coded
// files are reported as unsupported.
//
// Decoding steps: Huffman decode of each scan into per-component quantized
// coefficients (DC difference + AC run/size pairs in zig-zag order; the
// progressive scans add bands and bit planes to them), then per 8x8 block
// dequantization, a separable float IDCT
// rounded to 8 bits, chroma upsampling by linear ("triangle") interpolation
// at the JFIF sample positions, and the JFIF YCbCr -> RGB matrix.  Outputs of
// different decoders of the same file may differ by a level or two (IDCT and
// upsampling rounding are not specified by the standard).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "pt_scene.h"

namespace pupil {

namespace {

constexpr int kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                             12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                             35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                             58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huffman {
    // Canonical code: codes of length l are [first[l], first[l] + count[l]),
    // their symbols start at index[l] in `symbols`.
    int first[17] = {0}, count[17] = {0}, index[17] = {0};
    unsigned char symbols[256] = {0};
    bool defined = false;
};

struct Component {
    int id = 0, h = 1, v = 1, tq = 0, td = 0, ta = 0;
    int bw = 0, bh = 0;  // blocks per line / column (padded to whole MCUs)
    int cbw = 0, cbh = 0;  // blocks covering the component itself (non-interleaved scans)
    int dc_pred = 0;
    std::vector<int16_t> coef;         // bw x bh blocks of 64 quantized coefficients, natural order
    std::vector<unsigned char> plane;  // bw * 8 x bh * 8 samples
};

struct Reader {
    const std::vector<unsigned char> &f;
    size_t p = 0;
    uint32_t bits = 0;
    int nbits = 0;
    bool marker_hit = false;

    explicit Reader(const std::vector<unsigned char> &file) : f(file) {}

    // Entropy-coded bytes: 0xFF 0x00 is a data 0xFF; any other marker ends
    // the segment (zeros are fed from then on).
    int fill_byte() {
        if (marker_hit || p >= f.size()) return 0;
        unsigned char b = f[p];
        if (b == 0xFF) {
            const unsigned char n = p + 1 < f.size() ? f[p + 1] : 0;
            if (n == 0x00) {
                p += 2;
                return 0xFF;
            }
            marker_hit = true;
            return 0;
        }
        ++p;
        return b;
    }
    int bit() {
        if (nbits == 0) {
            bits = (uint32_t)fill_byte();
            nbits = 8;
        }
        --nbits;
        return (int)((bits >> nbits) & 1u);
    }
    int receive(int s) {
        int v = 0;
        for (int i = 0; i < s; ++i) v = (v << 1) | bit();
        return v;
    }
    bool decode(const Huffman &hf, int &sym) {
        int code = 0;
        for (int l = 1; l <= 16; ++l) {
            code = (code << 1) | bit();
            if (code - hf.first[l] < hf.count[l] && code >= hf.first[l]) {
                sym = hf.symbols[hf.index[l] + code - hf.first[l]];
                return true;
            }
        }
        return false;
    }
    // Restart: drop buffered bits, expect RSTn.
    bool restart() {
        nbits = 0;
        marker_hit = false;
        if (p + 1 < f.size() && f[p] == 0xFF && f[p + 1] >= 0xD0 && f[p + 1] <= 0xD7) {
            p += 2;
            return true;
        }
        return false;
    }
};

int extend(int v, int s) { return s == 0 ? 0 : (v < (1 << (s - 1)) ? v - (1 << s) + 1 : v); }

// Separable 8x8 inverse DCT (T.81 A.3.3) of dequantized coefficients in
// natural order; samples level-shifted by 128, rounded and clamped.
void idct8x8(const float in[64], unsigned char *out, int stride) {
    static float c[8][8];
    static bool init = false;
    if (!init) {
        for (int x = 0; x < 8; ++x)
            for (int u = 0; u < 8; ++u)
                c[x][u] = (float)((u == 0 ? std::sqrt(0.5) : 1.0) *
                                  std::cos((2 * x + 1) * u * 3.14159265358979323846 / 16.0) * 0.5);
        init = true;
    }
    float tmp[64];
    for (int y = 0; y < 8; ++y)  // rows of the output: sum over v
        for (int u = 0; u < 8; ++u) {
            float s = 0.f;
            for (int v = 0; v < 8; ++v) s += c[y][v] * in[v * 8 + u];
            tmp[y * 8 + u] = s;
        }
    for (int y = 0; y < 8; ++y)
        for (int x = 0; x < 8; ++x) {
            float s = 0.f;
            for (int u = 0; u < 8; ++u) s += c[x][u] * tmp[y * 8 + u];
            const int v = (int)std::lround(s + 128.f);
            out[y * stride + x] = (unsigned char)(v < 0 ? 0 : (v > 255 ? 255 : v));
        }
}

// Upsamples a component plane by (fx, fy) with linear interpolation between
// the sample centers (JFIF co-sited-at-center positions), edges clamped.
void upsample(const std::vector<unsigned char> &src, int sw, int sh, int fx, int fy, int ow, int oh,
              std::vector<unsigned char> &dst) {
    dst.resize((size_t)ow * oh);
    for (int y = 0; y < oh; ++y) {
        const float sy = ((float)y + 0.5f) / (float)fy - 0.5f;
        int y0 = (int)std::floor(sy);
        const float ty = sy - (float)y0;
        const int ya = std::min(std::max(y0, 0), sh - 1), yb = std::min(std::max(y0 + 1, 0), sh - 1);
        for (int x = 0; x < ow; ++x) {
            const float sx = ((float)x + 0.5f) / (float)fx - 0.5f;
            int x0 = (int)std::floor(sx);
            const float tx = sx - (float)x0;
            const int xa = std::min(std::max(x0, 0), sw - 1), xb = std::min(std::max(x0 + 1, 0), sw - 1);
            const float a = src[(size_t)ya * sw + xa] * (1.f - tx) + src[(size_t)ya * sw + xb] * tx;
            const float b = src[(size_t)yb * sw + xa] * (1.f - tx) + src[(size_t)yb * sw + xb] * tx;
            const int v = (int)std::lround(a * (1.f - ty) + b * ty);
            dst[(size_t)y * ow + x] = (unsigned char)(v < 0 ? 0 : (v > 255 ? 255 : v));
        }
    }
}

}  // namespace

// Decodes a JPEG file into 8-bit pixels (row 0 = top; 1 channel for gray
// images, 3 for color).
bool load_jpeg(const std::vector<unsigned char> &f, uint32_t &w, uint32_t &h, std::vector<unsigned char> &px,
               int &chans, std::string &err) {
    if (f.size() < 4 || f[0] != 0xFF || f[1] != 0xD8) {
        err = "parse: not a JPEG file";
        return false;
    }
    uint16_t qt[4][64];
    bool qt_defined[4] = {false, false, false, false};
    Huffman dc[4], ac[4];
    std::vector<Component> comp;
    int W = 0, H = 0, hmax = 1, vmax = 1, restart_interval = 0;
    bool frame = false, progressive = false, adobe_rgb = false;
    int scans = 0;
    size_t p = 2;
    auto seg_len = [&](size_t at) -> size_t { return (size_t)f[at] << 8 | f[at + 1]; };
    while (true) {
        while (p < f.size() && f[p] != 0xFF) ++p;  // skip fill up to a marker
        while (p < f.size() && f[p] == 0xFF) ++p;
        if (p >= f.size()) {
            if (scans > 0) break;  // image data without EOI: keep what was decoded
            err = "parse: JPEG ends before its image data";
            return false;
        }
        const unsigned char m = f[p++];
        if (m == 0xD9) break;                        // EOI
        if (m == 0x01 || (m >= 0xD0 && m <= 0xD7)) continue;
        if (p + 2 > f.size()) {
            err = "parse: truncated JPEG segment";
            return false;
        }
        const size_t len = seg_len(p);
        if (len < 2 || p + len > f.size()) {
            err = "parse: truncated JPEG segment";
            return false;
        }
        const unsigned char *d = &f[p + 2];
        const size_t n = len - 2;
        if (m == 0xDB) {  // DQT
            size_t q = 0;
            while (q < n) {
                const int pq = d[q] >> 4, tq = d[q] & 15;
                if (tq > 3 || q + 1 + 64 * (pq + 1) > n) {
                    err = "parse: bad JPEG quantization table";
                    return false;
                }
                for (int k = 0; k < 64; ++k)
                    qt[tq][kZigzag[k]] = pq ? (uint16_t)(d[q + 1 + 2 * k] << 8 | d[q + 2 + 2 * k]) : d[q + 1 + k];
                qt_defined[tq] = true;
                q += 1 + 64 * (pq + 1);
            }
        } else if (m == 0xC4) {  // DHT
            size_t q = 0;
            while (q < n) {
                if (q + 17 > n) {
                    err = "parse: bad JPEG Huffman table";
                    return false;
                }
                const int tc = d[q] >> 4, th = d[q] & 15;
                if (tc > 1 || th > 3) {
                    err = "parse: bad JPEG Huffman table";
                    return false;
                }
                Huffman &hf = tc ? ac[th] : dc[th];
                int total = 0, code = 0;
                for (int l = 1; l <= 16; ++l) {
                    hf.count[l] = d[q + l];
                    hf.first[l] = code;
                    hf.index[l] = total;
                    total += hf.count[l];
                    code = (code + hf.count[l]) << 1;
                }
                if (total > 256 || q + 17 + total > n) {
                    err = "parse: bad JPEG Huffman table";
                    return false;
                }
                std::memcpy(hf.symbols, d + q + 17, total);
                hf.defined = true;
                q += 17 + total;
            }
        } else if (m == 0xC0 || m == 0xC1 || m == 0xC2) {  // SOF0 / SOF1 / SOF2
            if (frame) {
                err = "parse: JPEG with two frame headers";
                return false;
            }
            if (n < 6 || d[0] != 8) {
                err = "unsupported: JPEG sample precision";
                return false;
            }
            progressive = m == 0xC2;
            H = d[1] << 8 | d[2];
            W = d[3] << 8 | d[4];
            const int nc = d[5];
            if (W <= 0 || H <= 0 || W > 65535 || H > 65535 || (nc != 1 && nc != 3) || n < 6 + 3 * (size_t)nc) {
                err = "unsupported: JPEG frame (1 or 3 components are read)";
                return false;
            }
            if (!image_size_ok((uint64_t)W, (uint64_t)H, err)) return false;
            comp.resize(nc);
            for (int k = 0; k < nc; ++k) {
                comp[k].id = d[6 + 3 * k];
                comp[k].h = d[7 + 3 * k] >> 4;
                comp[k].v = d[7 + 3 * k] & 15;
                comp[k].tq = d[8 + 3 * k];
                if (comp[k].h < 1 || comp[k].h > 4 || comp[k].v < 1 || comp[k].v > 4 || comp[k].tq > 3) {
                    err = "parse: bad JPEG component";
                    return false;
                }
                hmax = std::max(hmax, comp[k].h);
                vmax = std::max(vmax, comp[k].v);
            }
            // Coefficient storage padded to whole MCUs; a single-component
            // frame is never interleaved, its blocks cover the image itself.
            const int mcux = (W + 8 * hmax - 1) / (8 * hmax), mcuy = (H + 8 * vmax - 1) / (8 * vmax);
            for (auto &c : comp) {
                c.cbw = ((W * c.h + hmax - 1) / hmax + 7) / 8;
                c.cbh = ((H * c.v + vmax - 1) / vmax + 7) / 8;
                c.bw = nc == 1 ? c.cbw : mcux * c.h;
                c.bh = nc == 1 ? c.cbh : mcuy * c.v;
                c.coef.assign((size_t)c.bw * c.bh * 64, 0);
            }
            frame = true;
        } else if (m == 0xC3 || (m >= 0xC5 && m <= 0xCF && m != 0xC8 && m != 0xCC)) {
            err = "unsupported: lossless or arithmetic-coded JPEG";
            return false;
        } else if (m == 0xDD) {  // DRI
            if (n < 2) {
                err = "parse: bad JPEG restart interval";
                return false;
            }
            restart_interval = d[0] << 8 | d[1];
        } else if (m == 0xEE && n >= 12 && std::memcmp(d, "Adobe", 5) == 0) {
            adobe_rgb = d[11] == 0;  // transform 0: components are RGB, not YCbCr
        } else if (m == 0xDA) {  // SOS
            if (!frame) {
                err = "parse: JPEG scan before frame header";
                return false;
            }
            const int ns = n >= 1 ? d[0] : 0;
            if (ns < 1 || ns > (int)comp.size() || n < 1 + 2 * (size_t)ns + 3) {
                err = "parse: bad JPEG scan header";
                return false;
            }
            const int ss = d[1 + 2 * ns], se = d[2 + 2 * ns], ah = d[3 + 2 * ns] >> 4, al = d[3 + 2 * ns] & 15;
            // Scan kinds (T.81 G.1.1.1.1): sequential scans carry the whole
            // band 0..63 at full precision; progressive scans carry the DC
            // (Ss = 0 = Se, any components) or one AC band of one component.
            if (progressive ? (ss > se || se > 63 || (ss == 0 && se != 0) || (ss > 0 && ns != 1) || al > 13 ||
                               (ah != 0 && ah != al + 1))
                            : (ss != 0 || se != 63 || ah != 0 || al != 0)) {
                err = "parse: bad JPEG scan parameters";
                return false;
            }
            std::vector<Component *> sc;
            for (int k = 0; k < ns; ++k) {
                const int cid = d[1 + 2 * k];
                Component *c = nullptr;
                for (auto &cc : comp)
                    if (cc.id == cid) c = &cc;
                if (!c || std::find(sc.begin(), sc.end(), c) != sc.end()) {
                    err = "parse: JPEG scan names an unknown component";
                    return false;
                }
                c->td = d[2 + 2 * k] >> 4;
                c->ta = d[2 + 2 * k] & 15;
                const bool need_dc = ss == 0 && ah == 0, need_ac = se > 0;
                if (c->td > 3 || c->ta > 3 || (need_dc && !dc[c->td].defined) || (need_ac && !ac[c->ta].defined) ||
                    !qt_defined[c->tq]) {
                    err = "parse: JPEG scan without its tables";
                    return false;
                }
                sc.push_back(c);
            }
            p += len;
            for (auto *c : sc) c->dc_pred = 0;
            Reader r(f);
            r.p = p;
            int eobrun = 0;  // blocks left in an end-of-band run (progressive AC scans)
            const int p1 = 1 << al, m1 = -(1 << al);
            // Correction bit of an already nonzero coefficient (G.1.2.3).
            auto refine = [&](int16_t &c) {
                if (r.bit() && (c & p1) == 0) c = (int16_t)(c >= 0 ? c + p1 : c + m1);
            };
            auto decode_block = [&](Component &c, int bx, int by) -> bool {
                int16_t *co = &c.coef[((size_t)by * c.bw + bx) * 64];
                int sym;
                if (ss == 0) {  // DC
                    if (ah == 0) {
                        if (!r.decode(dc[c.td], sym) || sym > 11) return false;
                        c.dc_pred += extend(r.receive(sym), sym);
                        co[0] = (int16_t)(c.dc_pred * (1 << al));
                    } else if (r.bit()) {
                        co[0] = (int16_t)(co[0] | p1);
                    }
                    if (se == 0) return true;
                }
                const int k0 = std::max(ss, 1);
                if (!progressive || ah == 0) {  // AC band, first pass (or sequential)
                    if (eobrun > 0) {
                        --eobrun;
                        return true;
                    }
                    for (int k = k0; k <= se;) {
                        if (!r.decode(ac[c.ta], sym)) return false;
                        const int run = sym >> 4, size = sym & 15;
                        if (size == 0) {
                            if (run != 15) {  // end of band, here or for the next 2^run + bits - 1 blocks
                                if (progressive) eobrun = (1 << run) - 1 + (run ? r.receive(run) : 0);
                                break;
                            }
                            k += 16;
                            continue;
                        }
                        k += run;
                        if (k > se) return false;
                        co[kZigzag[k]] = (int16_t)(extend(r.receive(size), size) * (1 << al));
                        ++k;
                    }
                    return true;
                }
                // AC band, refinement pass: new coefficients of magnitude 1 << al,
                // each after `run` zero-history positions; every nonzero
                // coefficient passed on the way receives a correction bit.
                int k = k0;
                if (eobrun == 0) {
                    for (; k <= se; ++k) {
                        if (!r.decode(ac[c.ta], sym)) return false;
                        int run = sym >> 4, size = sym & 15, val = 0;
                        if (size != 0) {
                            if (size != 1) return false;
                            val = r.bit() ? p1 : m1;
                        } else if (run != 15) {
                            eobrun = (1 << run) + (run ? r.receive(run) : 0);
                            break;
                        }
                        for (; k <= se; ++k) {
                            int16_t &cz = co[kZigzag[k]];
                            if (cz != 0) refine(cz);
                            else if (run-- == 0) break;
                        }
                        if (val != 0) {
                            if (k > se) return false;
                            co[kZigzag[k]] = (int16_t)val;
                        }
                    }
                }
                if (eobrun > 0) {  // the rest of the band: correction bits only
                    for (; k <= se; ++k) {
                        int16_t &cz = co[kZigzag[k]];
                        if (cz != 0) refine(cz);
                    }
                    --eobrun;
                }
                return true;
            };
            const int mcux = (W + 8 * hmax - 1) / (8 * hmax), mcuy = (H + 8 * vmax - 1) / (8 * vmax);
            const bool interleaved = ns > 1;
            const int units = interleaved ? mcux * mcuy : sc[0]->cbw * sc[0]->cbh;
            for (int u = 0; u < units; ++u) {
                if (restart_interval && u > 0 && u % restart_interval == 0) {
                    if (!r.restart()) {
                        err = "parse: JPEG restart marker missing";
                        return false;
                    }
                    for (auto *c : sc) c->dc_pred = 0;
                    eobrun = 0;
                }
                bool ok = true;
                if (!interleaved) {
                    ok = decode_block(*sc[0], u % sc[0]->cbw, u / sc[0]->cbw);
                } else {
                    const int mx = u % mcux, my = u / mcux;
                    for (auto *c : sc)
                        for (int yy = 0; yy < c->v && ok; ++yy)
                            for (int xx = 0; xx < c->h && ok; ++xx)
                                ok = decode_block(*c, mx * c->h + xx, my * c->v + yy);
                }
                if (!ok) {
                    err = "parse: bad JPEG entropy-coded data";
                    return false;
                }
            }
            p = r.p;
            ++scans;
            continue;
        }
        p += len;
    }
    if (!frame || scans == 0) {
        err = "parse: JPEG without image data";
        return false;
    }
    // Dequantize and transform every block of every component.
    for (auto &c : comp) {
        c.plane.assign((size_t)c.bw * 8 * c.bh * 8, 0);
        const uint16_t *q = qt[c.tq];
        float blk[64];
        for (int by = 0; by < c.bh; ++by)
            for (int bx = 0; bx < c.bw; ++bx) {
                const int16_t *co = &c.coef[((size_t)by * c.bw + bx) * 64];
                for (int k = 0; k < 64; ++k) blk[k] = (float)((int)co[k] * (int)q[k]);
                idct8x8(blk, &c.plane[(size_t)by * 8 * c.bw * 8 + (size_t)bx * 8], c.bw * 8);
            }
        c.coef.clear();
        c.coef.shrink_to_fit();
    }
    w = (uint32_t)W;
    h = (uint32_t)H;
    chans = comp.size() == 1 ? 1 : 3;
    px.assign((size_t)W * H * chans, 0);
    std::vector<std::vector<unsigned char>> full(comp.size());
    for (size_t k = 0; k < comp.size(); ++k) {
        const Component &c = comp[k];
        const int fx = hmax / c.h, fy = vmax / c.v;
        if (comp.size() > 1 && (hmax % c.h != 0 || vmax % c.v != 0)) {
            err = "unsupported: JPEG sampling factors";
            return false;
        }
        const int cw = (W * c.h + hmax - 1) / hmax, ch = (H * c.v + vmax - 1) / vmax;  // component size
        std::vector<unsigned char> cp((size_t)cw * ch);
        for (int y = 0; y < ch; ++y)
            std::memcpy(&cp[(size_t)y * cw], &c.plane[(size_t)y * c.bw * 8], (size_t)cw);
        if (comp.size() == 1 || (fx == 1 && fy == 1)) full[k] = std::move(cp);
        else upsample(cp, cw, ch, fx, fy, W, H, full[k]);
    }
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        if (chans == 1) {
            px[i] = full[0][i];
            continue;
        }
        const float Y = full[0][i], Cb = full[1][i] - 128.f, Cr = full[2][i] - 128.f;
        float rgb[3];
        if (adobe_rgb) {
            rgb[0] = full[0][i];
            rgb[1] = full[1][i];
            rgb[2] = full[2][i];
        } else {
            rgb[0] = Y + 1.402f * Cr;
            rgb[1] = Y - 0.344136f * Cb - 0.714136f * Cr;
            rgb[2] = Y + 1.772f * Cb;
        }
        for (int k = 0; k < 3; ++k) {
            const int v = (int)std::lround(rgb[k]);
            px[3 * i + k] = (unsigned char)(v < 0 ? 0 : (v > 255 ? 255 : v));
        }
    }
    return true;
}

}  // namespace pupil
