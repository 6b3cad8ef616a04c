// image_io.cpp — see image_io.h. The decoders follow the PNG specification (ISO/IEC
// 15948) and the JPEG baseline process (ITU-T T.81 Annexes B, C, F): written here, with
// zlib only for PNG's deflate streams.
#include "image_io.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <stdexcept>

namespace msplat_hip {

namespace {

[[noreturn]] void bad(const std::string& what) { throw std::runtime_error(what); }

std::vector<uint8_t> read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) bad("cannot open image " + path);
    f.seekg(0, std::ios::end);
    const std::streamoff n = f.tellg();
    if (n <= 0) bad("empty image file " + path);
    f.seekg(0);
    std::vector<uint8_t> b((size_t)n);
    f.read(reinterpret_cast<char*>(b.data()), n);
    if (!f) bad("cannot read image " + path);
    return b;
}

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }
uint16_t be16(const uint8_t* p) { return (uint16_t)(p[0] << 8 | p[1]); }

// RGBA -> RGB over black: c * a / 255 to nearest
void put_rgba(uint8_t* out, uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
    if (a == 255) {
        out[0] = (uint8_t)r; out[1] = (uint8_t)g; out[2] = (uint8_t)b;
    } else {
        out[0] = (uint8_t)((r * a + 127) / 255);
        out[1] = (uint8_t)((g * a + 127) / 255);
        out[2] = (uint8_t)((b * a + 127) / 255);
    }
}

// ------------------------------------------------------------------------------------ PNG
struct PngHeader {
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = 0, interlace = 0;
    int channels() const { return ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4; }
    size_t row_bytes(uint32_t width) const { return ((size_t)width * channels() * depth + 7) / 8; }
    int filter_bpp() const { return std::max(1, channels() * depth / 8); }
};

uint8_t paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    return (uint8_t)(pb <= pc ? b : c);
}

// Reverses the per-row filters of one (sub-)image in place; rows are [filter byte | data].
void unfilter(uint8_t* data, size_t rows, size_t rb, int bpp) {
    std::vector<uint8_t> zero(rb, 0);
    const uint8_t* prev = zero.data();
    for (size_t y = 0; y < rows; y++) {
        uint8_t* row = data + y * (rb + 1);
        const int f = row[0];
        uint8_t* cur = row + 1;
        for (size_t i = 0; i < rb; i++) {
            const int a = i >= (size_t)bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= (size_t)bpp ? prev[i - bpp] : 0;
            switch (f) {
                case 0: break;
                case 1: cur[i] = (uint8_t)(cur[i] + a); break;
                case 2: cur[i] = (uint8_t)(cur[i] + b); break;
                case 3: cur[i] = (uint8_t)(cur[i] + ((a + b) >> 1)); break;
                case 4: cur[i] = (uint8_t)(cur[i] + paeth(a, b, c)); break;
                default: bad("PNG: bad filter type");
            }
        }
        prev = cur;
    }
}

uint32_t sample(const uint8_t* row, size_t idx, int depth) {
    if (depth == 8) return row[idx];
    if (depth == 16) return be16(row + 2 * idx);
    const size_t bit = idx * depth;
    const int shift = 8 - depth - (int)(bit & 7);
    return (row[bit >> 3] >> shift) & ((1u << depth) - 1);
}

}  // namespace

std::vector<uint8_t> decode_png(const uint8_t* d, size_t n, int& width, int& height) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (n < 8 || std::memcmp(d, sig, 8) != 0) bad("not a PNG file");
    PngHeader H;
    std::vector<uint8_t> idat, plte, trns;
    bool have_ihdr = false, end = false;
    size_t p = 8;
    while (p + 12 <= n && !end) {
        const uint32_t len = be32(d + p);
        if (len > n - p - 12) bad("PNG: truncated chunk");
        const uint8_t* type = d + p + 4;
        const uint8_t* body = d + p + 8;
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len < 13) bad("PNG: short IHDR");
            H.w = be32(body); H.h = be32(body + 4);
            H.depth = body[8]; H.ctype = body[9]; H.interlace = body[12];
            if (body[10] != 0 || body[11] != 0) bad("PNG: unknown compression or filter method");
            have_ihdr = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(body, body + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(body, body + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            end = true;
        }
        p += (size_t)len + 12;
    }
    if (!have_ihdr || H.w == 0 || H.h == 0 || H.w > (1u << 24) || H.h > (1u << 24)) bad("PNG: bad header");
    const bool ok_depth = (H.ctype == 0 && (H.depth == 1 || H.depth == 2 || H.depth == 4 || H.depth == 8 || H.depth == 16)) ||
                          (H.ctype == 3 && (H.depth == 1 || H.depth == 2 || H.depth == 4 || H.depth == 8)) ||
                          ((H.ctype == 2 || H.ctype == 4 || H.ctype == 6) && (H.depth == 8 || H.depth == 16));
    if (!ok_depth || H.interlace > 1) bad("PNG: unsupported colour type / bit depth / interlace");
    if (H.ctype == 3 && plte.size() < 3) bad("PNG: palette image without PLTE");
    // passes: Adam7 or one
    struct Pass { uint32_t x0, y0, dx, dy; };
    const Pass adam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4}, {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const Pass one[1] = {{0, 0, 1, 1}};
    const Pass* passes = H.interlace ? adam7 : one;
    const int np = H.interlace ? 7 : 1;
    size_t raw_size = 0;
    for (int k = 0; k < np; k++) {
        const uint32_t pw = (H.w - passes[k].x0 + passes[k].dx - 1) / passes[k].dx;
        const uint32_t ph = (H.h - passes[k].y0 + passes[k].dy - 1) / passes[k].dy;
        if (H.w > passes[k].x0 && H.h > passes[k].y0 && pw && ph) raw_size += (size_t)ph * (H.row_bytes(pw) + 1);
    }
    std::vector<uint8_t> raw(raw_size);
    {
        z_stream zs{};
        if (inflateInit(&zs) != Z_OK) bad("PNG: inflateInit failed");
        zs.next_in = idat.data();
        zs.avail_in = (uInt)idat.size();
        zs.next_out = raw.data();
        zs.avail_out = (uInt)raw.size();
        const int r = inflate(&zs, Z_FINISH);
        const size_t got = raw.size() - zs.avail_out;
        inflateEnd(&zs);
        if ((r != Z_STREAM_END && r != Z_BUF_ERROR && r != Z_OK) || got != raw.size()) bad("PNG: corrupt image data");
    }
    width = (int)H.w;
    height = (int)H.h;
    std::vector<uint8_t> out((size_t)H.w * H.h * 3);
    const int ch = H.channels();
    const uint32_t maxv = (1u << H.depth) - 1;
    auto to8 = [&](uint32_t v) -> uint32_t {
        if (H.depth == 8) return v;
        if (H.depth == 16) return (v * 255u + 32767u) / 65535u;
        return v * 255u / maxv;
    };
    // transparency key (grey or RGB), in the image's own sample depth
    const bool key_grey = H.ctype == 0 && trns.size() >= 2, key_rgb = H.ctype == 2 && trns.size() >= 6;
    size_t off = 0;
    for (int k = 0; k < np; k++) {
        const Pass& ps = passes[k];
        if (H.w <= ps.x0 || H.h <= ps.y0) continue;
        const uint32_t pw = (H.w - ps.x0 + ps.dx - 1) / ps.dx, ph = (H.h - ps.y0 + ps.dy - 1) / ps.dy;
        if (!pw || !ph) continue;
        const size_t rb = H.row_bytes(pw);
        uint8_t* sub = raw.data() + off;
        unfilter(sub, ph, rb, H.filter_bpp());
        for (uint32_t y = 0; y < ph; y++) {
            const uint8_t* row = sub + (size_t)y * (rb + 1) + 1;
            for (uint32_t x = 0; x < pw; x++) {
                uint32_t r, g, b, a = 255;
                const size_t i0 = (size_t)x * ch;
                if (H.ctype == 3) {
                    const uint32_t idx = sample(row, x, H.depth);
                    if ((size_t)idx * 3 + 2 >= plte.size()) bad("PNG: palette index out of range");
                    r = plte[3 * idx]; g = plte[3 * idx + 1]; b = plte[3 * idx + 2];
                    if (idx < trns.size()) a = trns[idx];
                } else if (H.ctype == 0 || H.ctype == 4) {
                    const uint32_t v = sample(row, i0, H.depth);
                    r = g = b = to8(v);
                    if (H.ctype == 4) a = to8(sample(row, i0 + 1, H.depth));
                    else if (key_grey && v == be16(trns.data())) a = 0;
                } else {
                    const uint32_t R = sample(row, i0, H.depth), G = sample(row, i0 + 1, H.depth),
                                   B = sample(row, i0 + 2, H.depth);
                    r = to8(R); g = to8(G); b = to8(B);
                    if (H.ctype == 6) a = to8(sample(row, i0 + 3, H.depth));
                    else if (key_rgb && R == be16(trns.data()) && G == be16(trns.data() + 2) && B == be16(trns.data() + 4))
                        a = 0;
                }
                const size_t X = ps.x0 + (size_t)x * ps.dx, Y = ps.y0 + (size_t)y * ps.dy;
                put_rgba(&out[(Y * H.w + X) * 3], r, g, b, a);
            }
        }
        off += (size_t)ph * (rb + 1);
    }
    return out;
}

// ----------------------------------------------------------------------------------- JPEG
namespace {

const int kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                         41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                         30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// Canonical Huffman decoding tables (T.81 Annex C and F.2.2.3): the largest code of each
// length, and where that length's values start.
struct Huff {
    bool defined = false;
    int32_t maxcode[18];
    int32_t valptr[17];
    int32_t mincode[17];
    uint8_t vals[256];
    // 9-bit lookahead: code length (0: longer) and value
    uint8_t look_len[512];
    uint8_t look_val[512];
    void build(const uint8_t counts[16], const uint8_t* symbols, int nsym) {
        std::memcpy(vals, symbols, (size_t)nsym);
        int code = 0, k = 0;
        std::memset(look_len, 0, sizeof(look_len));
        for (int l = 1; l <= 16; l++) {
            const int c = counts[l - 1];
            if (c == 0) {
                maxcode[l] = -1;
            } else {
                valptr[l] = k;
                mincode[l] = code;
                for (int i = 0; i < c; i++, k++, code++) {
                    if (l <= 9) {
                        const int lo = code << (9 - l), hi = (code + 1) << (9 - l);
                        for (int t = lo; t < hi; t++) {
                            look_len[t] = (uint8_t)l;
                            look_val[t] = vals[k];
                        }
                    }
                }
                maxcode[l] = code - 1;
            }
            code <<= 1;
        }
        maxcode[17] = 0x7FFFFFFF;
        defined = true;
    }
};

struct BitReader {
    const uint8_t* d;
    size_t n, p;
    uint32_t acc = 0;
    int bits = 0;
    bool marker_hit = false;
    int marker = 0;
    void fill() {
        while (bits <= 24) {
            uint32_t byte = 0;
            if (!marker_hit && p < n) {
                byte = d[p];
                if (byte == 0xFF) {
                    size_t q = p + 1;
                    while (q < n && d[q] == 0xFF) q++;  // fill bytes
                    const uint8_t m = q < n ? d[q] : 0xD9;
                    if (m == 0x00) {
                        p = q + 1;
                    } else {  // a marker ends the entropy-coded segment: feed zeros
                        marker_hit = true;
                        marker = m;
                        byte = 0;
                    }
                } else {
                    p++;
                }
            }
            acc |= byte << (24 - bits);
            bits += 8;
        }
    }
    uint32_t peek(int k) {
        if (bits < k) fill();
        return acc >> (32 - k);
    }
    void skip(int k) {
        acc <<= k;
        bits -= k;
    }
    uint32_t get(int k) {
        if (k == 0) return 0;
        const uint32_t v = peek(k);
        skip(k);
        return v;
    }
    // Restart marker: discard the partial byte, expect RSTn, resynchronise.
    void restart() {
        acc = 0;
        bits = 0;
        if (marker_hit) {
            if (marker < 0xD0 || marker > 0xD7) bad("JPEG: expected a restart marker");
            // step past FF.. RSTn
            while (p < n && d[p] == 0xFF) p++;
            p++;
            marker_hit = false;
        } else {
            while (p + 1 < n && !(d[p] == 0xFF && d[p + 1] >= 0xD0 && d[p + 1] <= 0xD7)) p++;
            if (p + 1 >= n) bad("JPEG: missing restart marker");
            p += 2;
        }
    }
};

int decode_huff(BitReader& br, const Huff& h) {
    const uint32_t look = br.peek(9);
    const int l9 = h.look_len[look];
    if (l9) {
        br.skip(l9);
        return h.look_val[look];
    }
    int l = 10;
    int32_t code = (int32_t)br.peek(10);
    while (l <= 16 && code > h.maxcode[l]) {
        l++;
        code = (int32_t)br.peek(l);
    }
    if (l > 16) bad("JPEG: bad Huffman code");
    br.skip(l);
    return h.vals[h.valptr[l] + code - h.mincode[l]];
}

int extend(uint32_t v, int s) { return s == 0 ? 0 : (v < (1u << (s - 1)) ? (int)v - (1 << s) + 1 : (int)v); }

struct Comp {
    int id = 0, h = 1, v = 1, tq = 0;
    int td = 0, ta = 0;
    int bw = 0, bh = 0;          // blocks across / down in the plane
    std::vector<uint8_t> plane;  // [bh * 8][bw * 8]
    int pred = 0;
};

// Separable 8x8 inverse DCT in double precision (T.81 A.3.3), level shift and clamp.
struct IdctTable {
    double c[8][8];  // c[x][u] = C(u) / 2 cos((2x + 1) u pi / 16)
    IdctTable() {
        const double pi = 3.14159265358979323846;
        for (int x = 0; x < 8; x++)
            for (int u = 0; u < 8; u++) c[x][u] = (u == 0 ? std::sqrt(0.5) : 1.0) * 0.5 * std::cos((2 * x + 1) * u * pi / 16.0);
    }
};

void idct_block(const int coef[64], uint8_t* dst, int stride) {
    static const IdctTable T;  // initialised once, thread-safe
    const auto& c = T.c;
    double tmp[64];
    for (int v = 0; v < 8; v++)
        for (int x = 0; x < 8; x++) {
            double s = 0.0;
            for (int u = 0; u < 8; u++) s += c[x][u] * coef[v * 8 + u];
            tmp[v * 8 + x] = s;
        }
    for (int y = 0; y < 8; y++)
        for (int x = 0; x < 8; x++) {
            double s = 0.0;
            for (int v = 0; v < 8; v++) s += c[y][v] * tmp[v * 8 + x];
            const long q = std::lround(s + 128.0);
            dst[y * stride + x] = (uint8_t)std::min(255L, std::max(0L, q));
        }
}

}  // namespace

std::vector<uint8_t> decode_jpeg(const uint8_t* d, size_t n, int& width, int& height) {
    if (n < 4 || d[0] != 0xFF || d[1] != 0xD8) bad("not a JPEG file");
    uint16_t qt[4][64];
    bool qdef[4] = {false, false, false, false};
    Huff hdc[4], hac[4];
    std::vector<Comp> comps;
    int W = 0, H = 0, hmax = 1, vmax = 1, restart_interval = 0, adobe_transform = -1;
    bool frame = false, done = false;
    size_t p = 2;
    while (p < n && !done) {
        if (d[p] != 0xFF) bad("JPEG: expected a marker");
        while (p < n && d[p] == 0xFF) p++;
        if (p >= n) break;
        const int m = d[p++];
        if (m == 0xD8 || (m >= 0xD0 && m <= 0xD7) || m == 0x01) continue;
        if (m == 0xD9) break;
        if (p + 2 > n) bad("JPEG: truncated segment");
        const size_t len = be16(d + p);
        if (len < 2 || p + len > n) bad("JPEG: bad segment length");
        const uint8_t* s = d + p + 2;
        const size_t sl = len - 2;
        switch (m) {
            case 0xDB: {  // DQT
                size_t q = 0;
                while (q < sl) {
                    const int pq = s[q] >> 4, tq = s[q] & 15;
                    if (tq > 3) bad("JPEG: bad quantisation table id");
                    q++;
                    for (int k = 0; k < 64; k++) {
                        if (pq) { qt[tq][k] = be16(s + q); q += 2; }
                        else { qt[tq][k] = s[q]; q += 1; }
                    }
                    qdef[tq] = true;
                }
                break;
            }
            case 0xC4: {  // DHT
                size_t q = 0;
                while (q + 17 <= sl) {
                    const int tc = s[q] >> 4, th = s[q] & 15;
                    if (th > 3 || tc > 1) bad("JPEG: bad Huffman table id");
                    int total = 0;
                    for (int k = 0; k < 16; k++) total += s[q + 1 + k];
                    if (total > 256 || q + 17 + total > sl) bad("JPEG: bad Huffman table");
                    (tc ? hac[th] : hdc[th]).build(s + q + 1, s + q + 17, total);
                    q += 17 + total;
                }
                break;
            }
            case 0xDD:  // DRI
                restart_interval = be16(s);
                break;
            case 0xEE:  // APP14 (Adobe): colour transform flag
                if (sl >= 12 && !std::memcmp(s, "Adobe", 5)) adobe_transform = s[11];
                break;
            case 0xC0:
            case 0xC1: {  // SOF0 / SOF1: sequential, Huffman
                if (s[0] != 8) bad("JPEG: only 8-bit samples are supported");
                H = be16(s + 1);
                W = be16(s + 3);
                const int nf = s[5];
                if (W <= 0 || H <= 0 || (nf != 1 && nf != 3)) bad("JPEG: unsupported frame (size or component count)");
                comps.resize(nf);
                for (int k = 0; k < nf; k++) {
                    comps[k].id = s[6 + 3 * k];
                    comps[k].h = s[7 + 3 * k] >> 4;
                    comps[k].v = s[7 + 3 * k] & 15;
                    comps[k].tq = s[8 + 3 * k] & 3;
                    if (comps[k].h < 1 || comps[k].h > 4 || comps[k].v < 1 || comps[k].v > 4) bad("JPEG: bad sampling");
                    hmax = std::max(hmax, comps[k].h);
                    vmax = std::max(vmax, comps[k].v);
                }
                const int mcux = (W + 8 * hmax - 1) / (8 * hmax), mcuy = (H + 8 * vmax - 1) / (8 * vmax);
                for (auto& c : comps) {
                    c.bw = mcux * c.h;
                    c.bh = mcuy * c.v;
                    c.plane.assign((size_t)c.bw * 8 * c.bh * 8, 0);
                }
                frame = true;
                break;
            }
            case 0xC2: case 0xC3: case 0xC5: case 0xC6: case 0xC7: case 0xC9: case 0xCA: case 0xCB: case 0xCD:
            case 0xCE: case 0xCF:
                bad("JPEG: progressive, lossless, hierarchical and arithmetic-coded files are not supported");
            case 0xDA: {  // SOS
                if (!frame) bad("JPEG: scan before frame");
                const int ns = s[0];
                if (ns < 1 || ns > (int)comps.size()) bad("JPEG: bad scan");
                std::vector<Comp*> sc;
                for (int k = 0; k < ns; k++) {
                    Comp* c = nullptr;
                    for (auto& cc : comps)
                        if (cc.id == s[1 + 2 * k]) c = &cc;
                    if (!c) bad("JPEG: scan names an unknown component");
                    c->td = s[2 + 2 * k] >> 4;
                    c->ta = s[2 + 2 * k] & 15;
                    if (c->td > 3 || c->ta > 3 || !hdc[c->td].defined || !hac[c->ta].defined || !qdef[c->tq])
                        bad("JPEG: scan uses an undefined table");
                    c->pred = 0;
                    sc.push_back(c);
                }
                BitReader br{d, n, p + len};
                int coef[64];
                auto block = [&](Comp& c, int bx, int by) {
                    int nat[64];
                    std::memset(nat, 0, sizeof(nat));
                    const int t = decode_huff(br, hdc[c.td]);
                    if (t > 11) bad("JPEG: bad DC size");
                    c.pred += extend(br.get(t), t);
                    nat[0] = c.pred * qt[c.tq][0];
                    for (int k = 1; k < 64;) {
                        const int rs = decode_huff(br, hac[c.ta]);
                        const int r = rs >> 4, sz = rs & 15;
                        if (sz == 0) {
                            if (r != 15) break;
                            k += 16;
                            continue;
                        }
                        k += r;
                        if (k > 63) bad("JPEG: coefficient index out of range");
                        nat[kZigzag[k]] = extend(br.get(sz), sz) * qt[c.tq][k];
                        k++;
                    }
                    std::memcpy(coef, nat, sizeof(coef));
                    const int stride = c.bw * 8;
                    idct_block(coef, &c.plane[(size_t)by * 8 * stride + (size_t)bx * 8], stride);
                };
                int mcus_x, mcus_y;
                if (ns == 1) {
                    Comp& c = *sc[0];
                    mcus_x = ((W * c.h + hmax - 1) / hmax + 7) / 8;
                    mcus_y = ((H * c.v + vmax - 1) / vmax + 7) / 8;
                } else {
                    mcus_x = (W + 8 * hmax - 1) / (8 * hmax);
                    mcus_y = (H + 8 * vmax - 1) / (8 * vmax);
                }
                int togo = restart_interval;
                for (int my = 0; my < mcus_y; my++)
                    for (int mx = 0; mx < mcus_x; mx++) {
                        if (restart_interval && togo == 0) {
                            br.restart();
                            for (Comp* c : sc) c->pred = 0;
                            togo = restart_interval;
                        }
                        if (ns == 1) {
                            block(*sc[0], mx, my);
                        } else {
                            for (Comp* c : sc)
                                for (int v = 0; v < c->v; v++)
                                    for (int h = 0; h < c->h; h++) block(*c, mx * c->h + h, my * c->v + v);
                        }
                        togo--;
                    }
                // continue after the entropy-coded data: at the marker that ended it
                size_t q = br.p;
                while (q + 1 < n && !(d[q] == 0xFF && d[q + 1] != 0x00 && !(d[q + 1] >= 0xD0 && d[q + 1] <= 0xD7))) q++;
                p = q;
                if (comps.size() == 1 || ns == (int)comps.size()) done = true;  // one scan holds every component
                continue;
            }
            default:
                break;  // APPn, COM and the rest: skipped
        }
        p += len;
    }
    if (!frame || comps.empty()) bad("JPEG: no frame");
    width = W;
    height = H;
    std::vector<uint8_t> out((size_t)W * H * 3);
    // each component sampled at full resolution: centred 2x upsampling with 3:1 weights (a
    // sample at full-resolution x lies at (x + 0.5) / 2 - 0.5 on its component's grid),
    // edges clamped; other ratios replicate
    std::vector<std::vector<float>> full(comps.size());
    for (size_t k = 0; k < comps.size(); k++) {
        const Comp& c = comps[k];
        const int rx = hmax / c.h, ry = vmax / c.v;
        const int cw = (W * c.h + hmax - 1) / hmax, chh = (H * c.v + vmax - 1) / vmax;
        const int stride = c.bw * 8;
        auto at = [&](int x, int y) -> float {
            x = std::min(std::max(x, 0), cw - 1);
            y = std::min(std::max(y, 0), chh - 1);
            return (float)c.plane[(size_t)y * stride + x];
        };
        std::vector<float>& f = full[k];
        f.resize((size_t)W * H);
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                float v;
                if (rx == 1 && ry == 1) {
                    v = at(x, y);
                } else if ((rx == 1 || rx == 2) && (ry == 1 || ry == 2)) {
                    int x0 = x, x1 = x, y0 = y, y1 = y;
                    float wx = 0.f, wy = 0.f;
                    if (rx == 2) { x0 = x >> 1; x1 = (x & 1) ? x0 + 1 : x0 - 1; wx = 0.25f; }
                    if (ry == 2) { y0 = y >> 1; y1 = (y & 1) ? y0 + 1 : y0 - 1; wy = 0.25f; }
                    v = (1.f - wx) * (1.f - wy) * at(x0, y0) + wx * (1.f - wy) * at(x1, y0) + (1.f - wx) * wy * at(x0, y1) +
                        wx * wy * at(x1, y1);
                } else {
                    v = at(x / rx, y / ry);
                }
                f[(size_t)y * W + x] = v;
            }
    }
    const bool rgb = comps.size() == 3 && (adobe_transform == 0 || (comps[0].id == 'R' && comps[1].id == 'G' && comps[2].id == 'B'));
    auto clamp8 = [](float v) -> uint8_t { return (uint8_t)std::min(255.f, std::max(0.f, std::nearbyint(v))); };
    for (size_t i = 0; i < (size_t)W * H; i++) {
        uint8_t* o = &out[3 * i];
        if (comps.size() == 1) {
            o[0] = o[1] = o[2] = clamp8(full[0][i]);
        } else if (rgb) {
            o[0] = clamp8(full[0][i]); o[1] = clamp8(full[1][i]); o[2] = clamp8(full[2][i]);
        } else {  // JFIF YCbCr (T.871)
            const float Y = full[0][i], cb = full[1][i] - 128.f, cr = full[2][i] - 128.f;
            o[0] = clamp8(Y + 1.402f * cr);
            o[1] = clamp8(Y - 0.344136f * cb - 0.714136f * cr);
            o[2] = clamp8(Y + 1.772f * cb);
        }
    }
    return out;
}

std::vector<uint8_t> decode_image_file(const std::string& path, int& width, int& height) {
    const std::vector<uint8_t> b = read_file(path);
    try {
        if (b.size() >= 8 && b[0] == 137 && b[1] == 'P' && b[2] == 'N' && b[3] == 'G') return decode_png(b.data(), b.size(), width, height);
        if (b.size() >= 3 && b[0] == 0xFF && b[1] == 0xD8) return decode_jpeg(b.data(), b.size(), width, height);
    } catch (const std::exception& e) {
        bad(path + ": " + e.what());
    }
    bad("unrecognised image format (PNG and JPEG are read): " + path);
}

HostImage read_image_rgb(const std::string& path) {
    HostImage img;
    const std::vector<uint8_t> b = decode_image_file(path, img.width, img.height);
    img.rgb.resize(b.size());
    for (size_t i = 0; i < b.size(); i++) img.rgb[i] = (float)b[i] / 255.0f;
    return img;
}

// Same float arithmetic as the reference's box filter: footprint [d*s, (d+1)*s) per axis,
// per-pixel weights (overlap widths, multiplied), sum of weighted samples times the
// reciprocal of the summed weight.
HostImage resize_area(const HostImage& src, int dw, int dh) {
    if (dw <= 0 || dh <= 0 || src.empty()) bad("resize_area: bad size");
    HostImage dst;
    dst.width = dw;
    dst.height = dh;
    dst.rgb.assign((size_t)dw * dh * 3, 0.f);
    const float sx = (float)src.width / (float)dw, sy = (float)src.height / (float)dh;
    for (int y = 0; y < dh; y++) {
        const float fy0 = (float)y * sy, fy1 = (float)(y + 1) * sy;
        const int iy0 = (int)fy0, iy1 = std::min((int)std::ceil(fy1), src.height);
        for (int x = 0; x < dw; x++) {
            const float fx0 = (float)x * sx, fx1 = (float)(x + 1) * sx;
            const int ix0 = (int)fx0, ix1 = std::min((int)std::ceil(fx1), src.width);
            float acc[3] = {0.f, 0.f, 0.f}, wsum = 0.f;
            for (int iy = iy0; iy < iy1; iy++) {
                const float wy = std::min((float)(iy + 1), fy1) - std::max((float)iy, fy0);
                for (int ix = ix0; ix < ix1; ix++) {
                    const float wx = std::min((float)(ix + 1), fx1) - std::max((float)ix, fx0);
                    const float w = wx * wy;
                    const float* s = &src.rgb[((size_t)iy * src.width + ix) * 3];
                    acc[0] += s[0] * w;
                    acc[1] += s[1] * w;
                    acc[2] += s[2] * w;
                    wsum += w;
                }
            }
            const float inv = 1.0f / wsum;
            float* o = &dst.rgb[((size_t)y * dw + x) * 3];
            o[0] = acc[0] * inv;
            o[1] = acc[1] * inv;
            o[2] = acc[2] * inv;
        }
    }
    return dst;
}

namespace {

// Brown-Conrady: undistorted normalised (x, y) -> distorted
void distort(float x, float y, float k1, float k2, float p1, float p2, float k3, float& xd, float& yd) {
    const float r2 = x * x + y * y, r4 = r2 * r2, r6 = r4 * r2;
    const float radial = 1.0f + k1 * r2 + k2 * r4 + k3 * r6;
    xd = x * radial + 2.0f * p1 * x * y + p2 * (r2 + 2.0f * x * x);
    yd = y * radial + p1 * (r2 + 2.0f * y * y) + 2.0f * p2 * x * y;
}

// the inverse by 20 fixed-point steps from the distorted point
void undistort_pt(float xd, float yd, float k1, float k2, float p1, float p2, float k3, float& xu, float& yu) {
    xu = xd;
    yu = yd;
    for (int it = 0; it < 20; it++) {
        const float r2 = xu * xu + yu * yu, r4 = r2 * r2, r6 = r4 * r2;
        const float radial = 1.0f + k1 * r2 + k2 * r4 + k3 * r6;
        const float dx = 2.0f * p1 * xu * yu + p2 * (r2 + 2.0f * xu * xu);
        const float dy = p1 * (r2 + 2.0f * yu * yu) + 2.0f * p2 * xu * yu;
        xu = (xd - dx) / radial;
        yu = (yd - dy) / radial;
    }
}

void bilinear(const HostImage& img, float x, float y, float out[3]) {
    const float fx = std::floor(x), fy = std::floor(y);
    const int xa = std::min(std::max((int)fx, 0), img.width - 1), xb = std::min(std::max((int)fx + 1, 0), img.width - 1);
    const int ya = std::min(std::max((int)fy, 0), img.height - 1), yb = std::min(std::max((int)fy + 1, 0), img.height - 1);
    const float tx = x - fx, ty = y - fy;
    const float* a = &img.rgb[((size_t)ya * img.width + xa) * 3];
    const float* b = &img.rgb[((size_t)ya * img.width + xb) * 3];
    const float* c = &img.rgb[((size_t)yb * img.width + xa) * 3];
    const float* e = &img.rgb[((size_t)yb * img.width + xb) * 3];
    for (int k = 0; k < 3; k++) {
        const float top = a[k] * (1.0f - tx) + b[k] * tx;
        const float bot = c[k] * (1.0f - tx) + e[k] * tx;
        out[k] = top * (1.0f - ty) + bot * ty;
    }
}

}  // namespace

Undistorted undistort_image(const HostImage& src, float fx, float fy, float cx, float cy, float k1, float k2, float p1,
                            float p2, float k3) {
    const int w = src.width, h = src.height;
    // inner edges of the undistorted border, from 200 samples per edge
    const int ns = 200;
    float top = -1e9f, bottom = 1e9f, left = -1e9f, right = 1e9f;
    for (int i = 0; i < ns; i++) {
        const float t = (float)i / (float)(ns - 1);
        float xu, yu;
        undistort_pt((t * (float)w - cx) / fx, (0.0f - cy) / fy, k1, k2, p1, p2, k3, xu, yu);
        top = std::max(top, yu * fy + cy);
        undistort_pt((t * (float)w - cx) / fx, ((float)(h - 1) - cy) / fy, k1, k2, p1, p2, k3, xu, yu);
        bottom = std::min(bottom, yu * fy + cy);
        undistort_pt((0.0f - cx) / fx, (t * (float)h - cy) / fy, k1, k2, p1, p2, k3, xu, yu);
        left = std::max(left, xu * fx + cx);
        undistort_pt(((float)(w - 1) - cx) / fx, (t * (float)h - cy) / fy, k1, k2, p1, p2, k3, xu, yu);
        right = std::min(right, xu * fx + cx);
    }
    int rx = std::max(0, (int)std::ceil(left)), ry = std::max(0, (int)std::ceil(top));
    int rw = std::min(w, (int)std::floor(right)) - rx, rh = std::min(h, (int)std::floor(bottom)) - ry;
    if (rw <= 0 || rh <= 0) {
        rx = 0; ry = 0; rw = w; rh = h;
    }
    Undistorted r;
    r.image.width = rw;
    r.image.height = rh;
    r.image.rgb.resize((size_t)rw * rh * 3);
    for (int oy = ry; oy < ry + rh; oy++)
        for (int ox = rx; ox < rx + rw; ox++) {
            float xd, yd;
            distort(((float)ox - cx) / fx, ((float)oy - cy) / fy, k1, k2, p1, p2, k3, xd, yd);
            bilinear(src, xd * fx + cx, yd * fy + cy, &r.image.rgb[((size_t)(oy - ry) * rw + (ox - rx)) * 3]);
        }
    r.fx = fx;
    r.fy = fy;
    r.cx = cx - (float)rx;
    r.cy = cy - (float)ry;
    return r;
}

std::vector<uint8_t> quantize_rgb(const HostImage& img) {
    std::vector<uint8_t> out(img.rgb.size());
    for (size_t i = 0; i < out.size(); i++)
        out[i] = (uint8_t)std::min(255.0f, std::max(0.0f, std::round(img.rgb[i] * 255.0f)));
    return out;
}

}  // namespace msplat_hip
