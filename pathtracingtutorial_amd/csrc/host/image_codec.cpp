// PNG / JPEG decoders and PNG / JPEG / BMP / TGA encoders (see image_codec.h).
#include "image_codec.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <fstream>

namespace ptamd {

bool read_file(const std::string& path, std::vector<uint8_t>& out, std::string* err) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) {
        if (err) *err = "cannot open '" + path + "'";
        return false;
    }
    f.seekg(0, std::ios::end);
    const std::streamoff n = f.tellg();
    f.seekg(0, std::ios::beg);
    out.resize(static_cast<size_t>(std::max<std::streamoff>(n, 0)));
    if (n > 0) f.read(reinterpret_cast<char*>(out.data()), n);
    if (!f) {
        if (err) *err = "cannot read '" + path + "'";
        return false;
    }
    return true;
}

namespace {

uint32_t be32(const uint8_t* p) {
    return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | uint32_t(p[3]);
}
uint16_t be16(const uint8_t* p) { return static_cast<uint16_t>((p[0] << 8) | p[1]); }

// ------------------------------------------------------------ CRC / Adler --
const std::array<uint32_t, 256>& crc_table() {
    static const std::array<uint32_t, 256> t = [] {
        std::array<uint32_t, 256> r{};
        for (uint32_t n = 0; n < 256; ++n) {
            uint32_t c = n;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;  // RFC 2083 15
            r[n] = c;
        }
        return r;
    }();
    return t;
}
uint32_t crc32(const uint8_t* p, size_t n, uint32_t c = 0xFFFFFFFFu) {
    const auto& t = crc_table();
    for (size_t i = 0; i < n; ++i) c = t[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
    return c;
}
uint32_t adler32(const uint8_t* p, size_t n) {
    uint32_t a = 1, b = 0;
    for (size_t i = 0; i < n; ++i) {
        a = (a + p[i]) % 65521u;
        b = (b + a) % 65521u;
    }
    return (b << 16) | a;
}

// ----------------------------------------------------------------- inflate --
// Bits are read least significant first (RFC 1951 3.1.1).
struct LsbBits {
    const uint8_t* p;
    size_t n, pos = 0;
    uint64_t buf = 0;
    int cnt = 0;
    bool over = false;
    LsbBits(const uint8_t* d, size_t len) : p(d), n(len) {}
    void fill() {
        while (cnt <= 56) {
            uint64_t byte = 0;
            if (pos < n) byte = p[pos];
            else if (pos >= n + 8) break;  // far past the end: stop padding
            ++pos;
            buf |= byte << cnt;
            cnt += 8;
        }
    }
    uint32_t peek(int k) {
        if (cnt < k) fill();
        return static_cast<uint32_t>(buf & ((1ull << k) - 1ull));
    }
    void drop(int k) {
        buf >>= k;
        cnt -= k;
        if (pos > n && static_cast<size_t>(cnt) < (pos - n) * 8) over = true;
    }
    uint32_t bits(int k) {
        if (k == 0) return 0;
        uint32_t v = peek(k);
        drop(k);
        return v;
    }
    void align() { drop(cnt & 7); }
};

// Canonical Huffman code (RFC 1951 3.2.2) with a 9-bit lookup of the
// bit-reversed short codes.
struct Huffman {
    uint16_t count[16] = {};
    uint16_t sym[320] = {};
    int16_t fast[512];
    bool build(const uint8_t* len, int n) {
        std::memset(count, 0, sizeof(count));
        for (int i = 0; i < n; ++i) count[len[i]]++;
        count[0] = 0;
        int left = 1;
        for (int l = 1; l < 16; ++l) {
            left = left * 2 - count[l];
            if (left < 0) return false;  // over-subscribed
        }
        uint16_t offs[16];
        offs[1] = 0;
        for (int l = 1; l < 15; ++l) offs[l + 1] = static_cast<uint16_t>(offs[l] + count[l]);
        for (int i = 0; i < n; ++i)
            if (len[i]) sym[offs[len[i]]++] = static_cast<uint16_t>(i);
        std::fill(std::begin(fast), std::end(fast), static_cast<int16_t>(-1));
        int code = 0, idx = 0;
        for (int l = 1; l <= 9; ++l) {
            for (int k = 0; k < count[l]; ++k, ++code, ++idx) {
                int rev = 0;
                for (int b = 0; b < l; ++b) rev |= ((code >> b) & 1) << (l - 1 - b);
                for (int r = rev; r < 512; r += 1 << l) fast[r] = static_cast<int16_t>((l << 9) | sym[idx]);
            }
            code <<= 1;
        }
        return true;
    }
    int decode(LsbBits& in) const {
        const uint32_t pk = in.peek(15);
        const int f = fast[pk & 511u];
        if (f >= 0) {
            in.drop(f >> 9);
            return f & 511;
        }
        int code = 0, first = 0, index = 0;
        for (int l = 1; l < 16; ++l) {
            code |= static_cast<int>((pk >> (l - 1)) & 1u);
            const int c = count[l];
            if (code - c < first) {
                in.drop(l);
                return sym[index + (code - first)];
            }
            index += c;
            first = (first + c) << 1;
            code <<= 1;
        }
        return -1;
    }
};

bool inflate_raw(const uint8_t* p, size_t n, std::vector<uint8_t>& out, std::string* err) {
    // length and distance bases (RFC 1951 3.2.5)
    int lbase[29], lextra[29], dbase[30], dextra[30];
    for (int i = 0; i < 8; ++i) {
        lbase[i] = 3 + i;
        lextra[i] = 0;
    }
    for (int i = 8; i < 28; ++i) {
        lextra[i] = (i - 4) / 4;
        lbase[i] = lbase[i - 1] + (1 << lextra[i - 1]);
    }
    lbase[28] = 258;
    lextra[28] = 0;
    for (int i = 0; i < 4; ++i) {
        dbase[i] = 1 + i;
        dextra[i] = 0;
    }
    for (int i = 4; i < 30; ++i) {
        dextra[i] = (i - 2) / 2;
        dbase[i] = dbase[i - 1] + (1 << dextra[i - 1]);
    }
    LsbBits in(p, n);
    Huffman lit, dist;
    bool final = false;
    while (!final) {
        final = in.bits(1) != 0;
        const uint32_t type = in.bits(2);
        if (type == 0) {
            in.align();
            const uint32_t len = in.bits(16), nlen = in.bits(16);
            if ((len ^ 0xFFFFu) != nlen) {
                if (err) *err = "deflate: bad stored block";
                return false;
            }
            for (uint32_t i = 0; i < len; ++i) out.push_back(static_cast<uint8_t>(in.bits(8)));
        } else if (type == 1 || type == 2) {
            uint8_t lens[320];
            int nlit = 288, ndist = 30;
            if (type == 1) {
                for (int i = 0; i < 144; ++i) lens[i] = 8;
                for (int i = 144; i < 256; ++i) lens[i] = 9;
                for (int i = 256; i < 280; ++i) lens[i] = 7;
                for (int i = 280; i < 288; ++i) lens[i] = 8;
                for (int i = 0; i < 30; ++i) lens[288 + i] = 5;
            } else {
                nlit = static_cast<int>(in.bits(5)) + 257;
                ndist = static_cast<int>(in.bits(5)) + 1;
                const int ncode = static_cast<int>(in.bits(4)) + 4;
                static const int order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
                uint8_t cl[19] = {};
                for (int i = 0; i < ncode; ++i) cl[order[i]] = static_cast<uint8_t>(in.bits(3));
                Huffman clh;
                if (!clh.build(cl, 19)) {
                    if (err) *err = "deflate: bad code-length code";
                    return false;
                }
                int i = 0;
                while (i < nlit + ndist) {
                    const int s = clh.decode(in);
                    if (s < 0) {
                        if (err) *err = "deflate: bad code length";
                        return false;
                    }
                    if (s < 16) {
                        lens[i++] = static_cast<uint8_t>(s);
                        continue;
                    }
                    int rep = 0;
                    uint8_t v = 0;
                    if (s == 16) {
                        if (i == 0) {
                            if (err) *err = "deflate: repeat with no previous length";
                            return false;
                        }
                        v = lens[i - 1];
                        rep = 3 + static_cast<int>(in.bits(2));
                    } else if (s == 17) {
                        rep = 3 + static_cast<int>(in.bits(3));
                    } else {
                        rep = 11 + static_cast<int>(in.bits(7));
                    }
                    if (i + rep > nlit + ndist) {
                        if (err) *err = "deflate: too many code lengths";
                        return false;
                    }
                    while (rep--) lens[i++] = v;
                }
                // lengths of the distance code follow the literal/length code
                std::memmove(lens + 288, lens + nlit, static_cast<size_t>(ndist));
            }
            if (!lit.build(lens, nlit) || !dist.build(lens + 288, ndist)) {
                if (err) *err = "deflate: bad Huffman code";
                return false;
            }
            while (true) {
                const int s = lit.decode(in);
                if (s < 0 || in.over) {
                    if (err) *err = "deflate: bad literal/length code";
                    return false;
                }
                if (s < 256) {
                    out.push_back(static_cast<uint8_t>(s));
                } else if (s == 256) {
                    break;
                } else {
                    const int li = s - 257;
                    if (li >= 29) {
                        if (err) *err = "deflate: bad length symbol";
                        return false;
                    }
                    const int len = lbase[li] + static_cast<int>(in.bits(lextra[li]));
                    const int ds = dist.decode(in);
                    if (ds < 0 || ds >= 30) {
                        if (err) *err = "deflate: bad distance symbol";
                        return false;
                    }
                    const size_t d = static_cast<size_t>(dbase[ds]) + in.bits(dextra[ds]);
                    if (d > out.size()) {
                        if (err) *err = "deflate: distance too far back";
                        return false;
                    }
                    const size_t from = out.size() - d;
                    for (int k = 0; k < len; ++k) out.push_back(out[from + k]);
                }
            }
        } else {
            if (err) *err = "deflate: bad block type";
            return false;
        }
        if (in.over) {
            if (err) *err = "deflate: truncated stream";
            return false;
        }
    }
    return true;
}

bool zlib_decode(const uint8_t* p, size_t n, std::vector<uint8_t>& out, std::string* err) {
    if (n < 6 || (p[0] & 0x0F) != 8 || ((p[0] << 8) | p[1]) % 31 != 0 || (p[1] & 0x20)) {
        if (err) *err = "zlib: bad header";
        return false;
    }
    return inflate_raw(p + 2, n - 2, out, err);
}

// --------------------------------------------------------------------- PNG --
int paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

}  // namespace

bool decode_png(const uint8_t* d, size_t n, Image8& img, std::string* err) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (n < 8 || std::memcmp(d, sig, 8) != 0) {
        if (err) *err = "not a PNG file";
        return false;
    }
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    size_t pos = 8;
    bool end = false;
    while (pos + 12 <= n && !end) {
        const uint32_t len = be32(d + pos);
        if (len > n - pos - 12) {
            if (err) *err = "PNG: truncated chunk";
            return false;
        }
        const uint8_t* type = d + pos + 4;
        const uint8_t* body = d + pos + 8;
        if ((crc32(type, len + 4) ^ 0xFFFFFFFFu) != be32(body + len)) {
            if (err) *err = "PNG: chunk CRC mismatch";
            return false;
        }
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len < 13) {
                if (err) *err = "PNG: bad IHDR";
                return false;
            }
            w = be32(body);
            h = be32(body + 4);
            depth = body[8];
            ctype = body[9];
            interlace = body[12];
            if (body[10] != 0 || body[11] != 0 || interlace > 1) {
                if (err) *err = "PNG: unsupported compression/filter/interlace method";
                return false;
            }
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(body, body + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(body, body + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            end = true;
        }
        pos += 12 + len;
    }
    static const int channels_of[7] = {1, 0, 3, 1, 2, 0, 4};
    const bool depth_ok = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                          (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                          ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (w == 0 || h == 0 || w > 65536 || h > 65536 || !depth_ok) {
        if (err) *err = "PNG: bad or unsupported header";
        return false;
    }
    if (ctype == 3 && plte.empty()) {
        if (err) *err = "PNG: palette image without PLTE";
        return false;
    }
    std::vector<uint8_t> raw;
    if (!zlib_decode(idat.data(), idat.size(), raw, err)) return false;
    const int ch = channels_of[ctype];
    const int bits_px = ch * depth;
    const int bpp = std::max(1, bits_px / 8);
    img.width = static_cast<int>(w);
    img.height = static_cast<int>(h);
    img.rgba.assign(static_cast<size_t>(w) * h * 4, 0);
    auto sample = [&](const uint8_t* row, uint32_t x, int c) -> uint32_t {
        if (depth == 8) return row[x * ch + c];
        if (depth == 16) return (uint32_t(row[(x * ch + c) * 2]) << 8) | row[(x * ch + c) * 2 + 1];
        const uint32_t bit = x * static_cast<uint32_t>(depth);  // depth < 8: one channel
        return (row[bit >> 3] >> (8 - depth - (bit & 7))) & ((1u << depth) - 1u);
    };
    auto to8 = [&](uint32_t v) -> uint8_t {
        if (depth == 16) return static_cast<uint8_t>(v >> 8);
        if (depth == 8) return static_cast<uint8_t>(v);
        return static_cast<uint8_t>(v * 255u / ((1u << depth) - 1u));
    };
    struct Pass { int x0, y0, dx, dy; };
    const Pass adam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4}, {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const Pass whole[1] = {{0, 0, 1, 1}};
    const Pass* passes = interlace ? adam7 : whole;
    const int npass = interlace ? 7 : 1;
    size_t off = 0;
    for (int ps = 0; ps < npass; ++ps) {
        const Pass& P = passes[ps];
        const uint32_t pw = (w > static_cast<uint32_t>(P.x0)) ? (w - P.x0 + P.dx - 1) / P.dx : 0;
        const uint32_t ph = (h > static_cast<uint32_t>(P.y0)) ? (h - P.y0 + P.dy - 1) / P.dy : 0;
        if (pw == 0 || ph == 0) continue;
        const size_t stride = (static_cast<size_t>(pw) * bits_px + 7) / 8;
        std::vector<uint8_t> prev(stride, 0), cur(stride);
        for (uint32_t y = 0; y < ph; ++y) {
            if (off + 1 + stride > raw.size()) {
                if (err) *err = "PNG: image data too short";
                return false;
            }
            const int f = raw[off];
            const uint8_t* src = raw.data() + off + 1;
            off += 1 + stride;
            for (size_t i = 0; i < stride; ++i) {
                const int a = i >= static_cast<size_t>(bpp) ? cur[i - bpp] : 0;
                const int b = prev[i];
                const int c = i >= static_cast<size_t>(bpp) ? prev[i - bpp] : 0;
                int v = src[i];
                switch (f) {
                    case 0: break;
                    case 1: v += a; break;
                    case 2: v += b; break;
                    case 3: v += (a + b) >> 1; break;
                    case 4: v += paeth(a, b, c); break;
                    default:
                        if (err) *err = "PNG: bad filter type";
                        return false;
                }
                cur[i] = static_cast<uint8_t>(v);
            }
            const uint32_t yy = P.y0 + y * P.dy;
            for (uint32_t x = 0; x < pw; ++x) {
                const uint32_t xx = P.x0 + x * P.dx;
                uint8_t* o = &img.rgba[(static_cast<size_t>(yy) * w + xx) * 4];
                if (ctype == 3) {
                    const uint32_t idx = sample(cur.data(), x, 0);
                    if (idx * 3 + 2 >= plte.size()) {
                        if (err) *err = "PNG: palette index out of range";
                        return false;
                    }
                    o[0] = plte[idx * 3];
                    o[1] = plte[idx * 3 + 1];
                    o[2] = plte[idx * 3 + 2];
                    o[3] = idx < trns.size() ? trns[idx] : 255;
                } else if (ctype == 0 || ctype == 4) {
                    const uint32_t g = sample(cur.data(), x, 0);
                    o[0] = o[1] = o[2] = to8(g);
                    if (ctype == 4) o[3] = to8(sample(cur.data(), x, 1));
                    else o[3] = (trns.size() >= 2 && g == be16(trns.data())) ? 0 : 255;
                } else {
                    const uint32_t r = sample(cur.data(), x, 0), g = sample(cur.data(), x, 1),
                                   b = sample(cur.data(), x, 2);
                    o[0] = to8(r);
                    o[1] = to8(g);
                    o[2] = to8(b);
                    if (ctype == 6) o[3] = to8(sample(cur.data(), x, 3));
                    else
                        o[3] = (trns.size() >= 6 && r == be16(trns.data()) && g == be16(trns.data() + 2) &&
                                b == be16(trns.data() + 4))
                                   ? 0
                                   : 255;
                }
            }
            std::swap(prev, cur);
        }
    }
    return true;
}

// -------------------------------------------------------------------- JPEG --
namespace {

// zigzag[k] = natural (row-major) index of the k-th coefficient in zigzag order
const std::array<int, 64>& zigzag() {
    static const std::array<int, 64> z = [] {
        std::array<int, 64> r{};
        int k = 0;
        for (int s = 0; s < 15; ++s) {  // anti-diagonals x + y = s, alternating direction
            if (s % 2 == 0) {
                for (int y = std::min(s, 7); y >= 0 && s - y <= 7; --y) r[k++] = y * 8 + (s - y);
            } else {
                for (int x = std::min(s, 7); x >= 0 && s - x <= 7; --x) r[k++] = (s - x) * 8 + x;
            }
        }
        return r;
    }();
    return z;
}

struct JHuff {
    bool present = false;
    uint8_t bits[17] = {};
    uint8_t vals[256] = {};
    int maxcode[18];
    int valptr[17];
    int mincode[17];
    uint16_t fast[512];  // (len << 8) | value for codes of <= 9 bits, 0 = slow path
    void build() {
        int code = 0, k = 0;
        std::memset(fast, 0, sizeof(fast));
        for (int l = 1; l <= 16; ++l) {
            valptr[l] = k;
            mincode[l] = code;
            for (int i = 0; i < bits[l]; ++i, ++k, ++code)
                if (l <= 9)
                    for (int r = code << (9 - l); r < ((code + 1) << (9 - l)); ++r)
                        fast[r] = static_cast<uint16_t>((l << 8) | vals[k]);
            maxcode[l] = bits[l] ? code - 1 : -1;
            code <<= 1;
        }
        maxcode[17] = 0x7FFFFFFF;
    }
};

// Entropy-coded segment reader: most significant bit first, 0xFF00 stuffing;
// a marker ends the data (zeros are read past it).
struct MsbBits {
    const uint8_t* p;
    size_t n, pos;
    uint32_t buf = 0;
    int cnt = 0;
    bool marker = false;
    MsbBits(const uint8_t* d, size_t len, size_t start) : p(d), n(len), pos(start) {}
    void fill() {
        while (cnt <= 24) {
            uint32_t b = 0;
            if (!marker && pos < n) {
                b = p[pos];
                if (b == 0xFF) {
                    const uint8_t nx = pos + 1 < n ? p[pos + 1] : 0;
                    if (nx == 0x00) {
                        pos += 2;
                    } else {
                        marker = true;
                        b = 0;
                    }
                } else {
                    ++pos;
                }
            }
            buf |= b << (24 - cnt);
            cnt += 8;
        }
    }
    uint32_t peek(int k) {
        if (cnt < k) fill();
        return buf >> (32 - k);
    }
    void drop(int k) {
        buf <<= k;
        cnt -= k;
    }
    uint32_t bits(int k) {
        if (k == 0) return 0;
        uint32_t v = peek(k);
        drop(k);
        return v;
    }
    int decode(const JHuff& h) {
        const uint32_t pk = peek(16);
        const uint16_t f = h.fast[pk >> 7];
        if (f) {
            drop(f >> 8);
            return f & 0xFF;
        }
        for (int l = 10; l <= 16; ++l) {
            const int code = static_cast<int>(pk >> (16 - l));
            if (code <= h.maxcode[l]) {
                drop(l);
                return h.vals[h.valptr[l] + code - h.mincode[l]];
            }
        }
        return -1;
    }
    // Skips to just past the next RSTn marker.
    void restart() {
        buf = 0;
        cnt = 0;
        marker = false;
        while (pos + 1 < n && !(p[pos] == 0xFF && p[pos + 1] >= 0xD0 && p[pos + 1] <= 0xD7)) ++pos;
        if (pos + 1 < n) pos += 2;
    }
};

int extend(int v, int s) { return (s == 0) ? 0 : (v < (1 << (s - 1)) ? v - (1 << s) + 1 : v); }

struct JComp {
    int id = 0, h = 1, v = 1, tq = 0, td = 0, ta = 0;
    int bw = 0, bh = 0;             // blocks per line / column covered by MCUs
    std::vector<int16_t> coef;      // bw * bh * 64, natural order
    int pred = 0;
    std::vector<uint8_t> plane;     // bw*8 x bh*8 samples
};

}  // namespace

bool decode_jpeg(const uint8_t* d, size_t n, Image8& img, std::string* err) {
    auto fail = [&](const char* m) {
        if (err) *err = std::string("JPEG: ") + m;
        return false;
    };
    if (n < 4 || d[0] != 0xFF || d[1] != 0xD8) return fail("not a JPEG file");
    const auto& zz = zigzag();
    uint16_t qt[4][64] = {};
    JHuff hdc[4], hac[4];
    std::vector<JComp> comps;
    int W = 0, H = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
    bool progressive = false, frame = false, adobe_rgb = false;
    int restart_interval = 0;
    int eobrun = 0;
    size_t pos = 2;
    while (pos + 4 <= n) {
        if (d[pos] != 0xFF) {
            ++pos;
            continue;
        }
        const uint8_t m = d[pos + 1];
        if (m == 0xFF) {
            ++pos;
            continue;
        }
        if (m == 0xD9) break;                          // EOI
        if (m == 0xD8 || (m >= 0xD0 && m <= 0xD7)) {  // stray SOI / RSTn
            pos += 2;
            continue;
        }
        const size_t len = be16(d + pos + 2);
        if (len < 2 || pos + 2 + len > n) return fail("truncated segment");
        const uint8_t* s = d + pos + 4;
        const size_t sl = len - 2;
        if (m == 0xDB) {  // DQT
            size_t i = 0;
            while (i < sl) {
                const int pq = s[i] >> 4, tq = s[i] & 15;
                if (tq > 3 || i + 1 + 64 * (pq + 1) > sl) return fail("bad DQT");
                for (int k = 0; k < 64; ++k)
                    qt[tq][k] = pq ? be16(s + i + 1 + 2 * k) : s[i + 1 + k];  // zigzag order
                i += 1 + 64 * (pq + 1);
            }
        } else if (m == 0xC4) {  // DHT
            size_t i = 0;
            while (i < sl) {
                const int tc = s[i] >> 4, th = s[i] & 15;
                if (tc > 1 || th > 3 || i + 17 > sl) return fail("bad DHT");
                JHuff& h = tc ? hac[th] : hdc[th];
                int total = 0;
                for (int l = 1; l <= 16; ++l) {
                    h.bits[l] = s[i + l];
                    total += h.bits[l];
                }
                if (total > 256 || i + 17 + total > sl) return fail("bad DHT");
                std::memcpy(h.vals, s + i + 17, static_cast<size_t>(total));
                h.build();
                h.present = true;
                i += 17 + total;
            }
        } else if (m == 0xC0 || m == 0xC1 || m == 0xC2) {  // SOF0/1/2
            if (sl < 6 || s[0] != 8) return fail("only 8-bit samples are supported");
            progressive = m == 0xC2;
            H = be16(s + 1);
            W = be16(s + 3);
            const int nc = s[5];
            if (W <= 0 || H <= 0 || (nc != 1 && nc != 3) || sl < 6u + 3u * nc) return fail("bad or unsupported frame");
            comps.assign(static_cast<size_t>(nc), JComp{});
            for (int c = 0; c < nc; ++c) {
                comps[c].id = s[6 + 3 * c];
                comps[c].h = s[7 + 3 * c] >> 4;
                comps[c].v = s[7 + 3 * c] & 15;
                comps[c].tq = s[8 + 3 * c] & 3;
                if (comps[c].h < 1 || comps[c].h > 4 || comps[c].v < 1 || comps[c].v > 4)
                    return fail("bad sampling factors");
                hmax = std::max(hmax, comps[c].h);
                vmax = std::max(vmax, comps[c].v);
            }
            mcux = (W + 8 * hmax - 1) / (8 * hmax);
            mcuy = (H + 8 * vmax - 1) / (8 * vmax);
            for (auto& c : comps) {
                c.bw = mcux * c.h;
                c.bh = mcuy * c.v;
                c.coef.assign(static_cast<size_t>(c.bw) * c.bh * 64, 0);
            }
            frame = true;
        } else if ((m >= 0xC3 && m <= 0xCF) && m != 0xC4 && m != 0xC8 && m != 0xCC) {
            return fail("lossless, hierarchical and arithmetic-coded JPEG are not supported");
        } else if (m == 0xDD) {  // DRI
            if (sl < 2) return fail("bad DRI");
            restart_interval = be16(s);
        } else if (m == 0xEE) {  // APP14: Adobe colour transform flag
            if (sl >= 12 && !std::memcmp(s, "Adobe", 5)) adobe_rgb = s[11] == 0;
        } else if (m == 0xDA) {  // SOS
            if (!frame) return fail("scan before frame");
            const int ns = s[0];
            if (ns < 1 || ns > 4 || sl < 4u + 2u * ns) return fail("bad SOS");
            std::vector<JComp*> sc;
            for (int i = 0; i < ns; ++i) {
                const int cid = s[1 + 2 * i];
                JComp* c = nullptr;
                for (auto& x : comps)
                    if (x.id == cid) c = &x;
                if (!c) return fail("scan names an unknown component");
                c->td = s[2 + 2 * i] >> 4;
                c->ta = s[2 + 2 * i] & 15;
                if (c->td > 3 || c->ta > 3) return fail("bad table selector");
                sc.push_back(c);
            }
            const int Ss = s[1 + 2 * ns], Se = s[2 + 2 * ns];
            const int Ah = s[3 + 2 * ns] >> 4, Al = s[3 + 2 * ns] & 15;
            if (!progressive && (Ss != 0 || Se != 63)) return fail("bad spectral selection");
            if (Ss > Se || Se > 63) return fail("bad spectral selection");
            for (JComp* c : sc) {
                if ((Ss == 0 && !hdc[c->td].present && Ah == 0) || (Se > 0 && Ss > 0 && !hac[c->ta].present) ||
                    (!progressive && !hac[c->ta].present))
                    return fail("scan uses an undefined Huffman table");
                c->pred = 0;
            }
            eobrun = 0;
            MsbBits in(d, n, pos + 2 + len);
            auto decode_block = [&](JComp& c, int16_t* blk) -> bool {
                if (!progressive) {
                    const int t = in.decode(hdc[c.td]);
                    if (t < 0 || t > 16) return false;
                    c.pred += extend(static_cast<int>(in.bits(t)), t);
                    blk[0] = static_cast<int16_t>(c.pred);
                    for (int k = 1; k < 64;) {
                        const int rs = in.decode(hac[c.ta]);
                        if (rs < 0) return false;
                        const int r = rs >> 4, sz = rs & 15;
                        if (sz == 0) {
                            if (r != 15) break;
                            k += 16;
                            continue;
                        }
                        k += r;
                        if (k > 63) return false;
                        blk[zz[k]] = static_cast<int16_t>(extend(static_cast<int>(in.bits(sz)), sz));
                        ++k;
                    }
                    return true;
                }
                if (Ss == 0) {  // DC scans (ITU T.81 G.1.2.1)
                    if (Ah == 0) {
                        const int t = in.decode(hdc[c.td]);
                        if (t < 0 || t > 16) return false;
                        c.pred += extend(static_cast<int>(in.bits(t)), t);
                        blk[0] = static_cast<int16_t>(c.pred * (1 << Al));
                    } else if (in.bits(1)) {
                        blk[0] = static_cast<int16_t>(blk[0] | (1 << Al));
                    }
                    return true;
                }
                if (Ah == 0) {  // AC first pass (G.1.2.2)
                    if (eobrun > 0) {
                        --eobrun;
                        return true;
                    }
                    for (int k = Ss; k <= Se;) {
                        const int rs = in.decode(hac[c.ta]);
                        if (rs < 0) return false;
                        const int r = rs >> 4, sz = rs & 15;
                        if (sz == 0) {
                            if (r < 15) {
                                eobrun = (1 << r) - 1;
                                if (r) eobrun += static_cast<int>(in.bits(r));
                                break;
                            }
                            k += 16;
                            continue;
                        }
                        k += r;
                        if (k > 63) return false;
                        blk[zz[k]] = static_cast<int16_t>(extend(static_cast<int>(in.bits(sz)), sz) * (1 << Al));
                        ++k;
                    }
                    return true;
                }
                // AC refinement (G.1.2.3): one more bit of each nonzero
                // coefficient in the band, and new coefficients of magnitude 1
                const int p1 = 1 << Al, m1 = -(1 << Al);
                int k = Ss;
                auto refine = [&](int idx) {
                    int16_t& v = blk[zz[idx]];
                    if (in.bits(1) && (v & p1) == 0) v = static_cast<int16_t>(v >= 0 ? v + p1 : v + m1);
                };
                if (eobrun == 0) {
                    for (; k <= Se;) {
                        const int rs = in.decode(hac[c.ta]);
                        if (rs < 0) return false;
                        int r = rs >> 4, sz = rs & 15;
                        int val = 0;
                        if (sz) {
                            if (sz != 1) return false;
                            val = in.bits(1) ? p1 : m1;
                        } else if (r != 15) {
                            eobrun = 1 << r;
                            if (r) eobrun += static_cast<int>(in.bits(r));
                            break;
                        }
                        while (k <= Se) {
                            if (blk[zz[k]] != 0) {
                                refine(k);
                            } else {
                                if (r == 0) {
                                    if (val) blk[zz[k]] = static_cast<int16_t>(val);
                                    ++k;
                                    break;
                                }
                                --r;
                            }
                            ++k;
                        }
                    }
                }
                if (eobrun > 0) {
                    for (; k <= Se; ++k)
                        if (blk[zz[k]] != 0) refine(k);
                    --eobrun;
                }
                return true;
            };
            auto block_at = [&](JComp& c, int bx, int by) { return c.coef.data() + (static_cast<size_t>(by) * c.bw + bx) * 64; };
            int mcus, mx_count;
            const bool interleaved = ns > 1;
            if (interleaved) {
                mx_count = mcux;
                mcus = mcux * mcuy;
            } else {  // a single component: one block per MCU over the component's own extent
                JComp& c = *sc[0];
                const int cw = (W * c.h + hmax - 1) / hmax, ch = (H * c.v + vmax - 1) / vmax;
                mx_count = (cw + 7) / 8;
                mcus = mx_count * ((ch + 7) / 8);
            }
            int todo = restart_interval;
            for (int mi = 0; mi < mcus; ++mi) {
                if (restart_interval && todo == 0) {
                    in.restart();
                    for (JComp* c : sc) c->pred = 0;
                    eobrun = 0;
                    todo = restart_interval;
                }
                const int mx = mi % mx_count, my = mi / mx_count;
                if (interleaved) {
                    for (JComp* c : sc)
                        for (int v = 0; v < c->v; ++v)
                            for (int h = 0; h < c->h; ++h)
                                if (!decode_block(*c, block_at(*c, mx * c->h + h, my * c->v + v)))
                                    return fail("corrupt entropy-coded data");
                } else {
                    if (!decode_block(*sc[0], block_at(*sc[0], mx, my))) return fail("corrupt entropy-coded data");
                }
                if (restart_interval) --todo;
            }
            // continue after the entropy-coded data: find the next marker
            size_t q = in.pos;
            while (q + 1 < n && !(d[q] == 0xFF && d[q + 1] != 0x00 && !(d[q + 1] >= 0xD0 && d[q + 1] <= 0xD7))) ++q;
            pos = q;
            continue;
        }
        pos += 2 + len;
    }
    if (!frame) return fail("no frame");
    // dequantize + inverse DCT (T.81 A.3.3) per block
    static const std::array<float, 64> cosines = [] {
        std::array<float, 64> t{};
        for (int x = 0; x < 8; ++x)
            for (int u = 0; u < 8; ++u)
                t[x * 8 + u] = static_cast<float>((u == 0 ? std::sqrt(0.5) : 1.0) *
                                                  std::cos((2 * x + 1) * u * 3.14159265358979323846 / 16.0) * 0.5);
        return t;
    }();
    for (auto& c : comps) {
        c.plane.assign(static_cast<size_t>(c.bw) * 8 * c.bh * 8, 0);
        const uint16_t* q = qt[c.tq];
        float F[64], tmp[64];
        for (int by = 0; by < c.bh; ++by)
            for (int bx = 0; bx < c.bw; ++bx) {
                const int16_t* blk = c.coef.data() + (static_cast<size_t>(by) * c.bw + bx) * 64;
                for (int k = 0; k < 64; ++k) F[zz[k]] = static_cast<float>(blk[zz[k]]) * q[k];
                for (int y = 0; y < 8; ++y)  // rows: sum over u
                    for (int x = 0; x < 8; ++x) {
                        float s = 0.0f;
                        for (int u = 0; u < 8; ++u) s += cosines[x * 8 + u] * F[y * 8 + u];
                        tmp[y * 8 + x] = s;
                    }
                for (int x = 0; x < 8; ++x)  // columns: sum over v
                    for (int y = 0; y < 8; ++y) {
                        float s = 0.0f;
                        for (int v = 0; v < 8; ++v) s += cosines[y * 8 + v] * tmp[v * 8 + x];
                        const int val = static_cast<int>(std::lround(s + 128.0f));
                        c.plane[(static_cast<size_t>(by) * 8 + y) * (c.bw * 8) + bx * 8 + x] =
                            static_cast<uint8_t>(std::clamp(val, 0, 255));
                    }
            }
    }
    // upsample to the full grid (triangle filter for 2x, replication otherwise) and convert
    auto sample_at = [&](const JComp& c, int x, int y) -> int {
        const int pw = c.bw * 8;
        const int fx = hmax / c.h, fy = vmax / c.v;
        if (fx == 1 && fy == 1) return c.plane[static_cast<size_t>(y) * pw + x];
        const int cw = (W * c.h + hmax - 1) / hmax, ch = (H * c.v + vmax - 1) / vmax;
        auto at = [&](int xx, int yy) {
            xx = std::clamp(xx, 0, cw - 1);
            yy = std::clamp(yy, 0, ch - 1);
            return static_cast<int>(c.plane[static_cast<size_t>(yy) * pw + xx]);
        };
        if ((fx == 2 || fx == 1) && (fy == 2 || fy == 1) && hmax % c.h == 0 && vmax % c.v == 0) {
            const int sx = x / fx, sy = y / fy;
            // vertical: 3/4 nearest row + 1/4 the row on the sample's side
            auto col = [&](int xx) {
                if (fy == 1) return 4 * at(xx, sy);
                const int other = (y % 2 == 0) ? sy - 1 : sy + 1;
                return 3 * at(xx, sy) + at(xx, other);
            };
            if (fx == 1) return (col(sx) + 2) >> 2;
            const int other = (x % 2 == 0) ? sx - 1 : sx + 1;
            const int v = 3 * col(sx) + col(other);
            return (v + 8) >> 4;
        }
        return at(x / fx, y / fy);
    };
    img.width = W;
    img.height = H;
    img.rgba.assign(static_cast<size_t>(W) * H * 4, 255);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            uint8_t* o = &img.rgba[(static_cast<size_t>(y) * W + x) * 4];
            if (comps.size() == 1) {
                o[0] = o[1] = o[2] = static_cast<uint8_t>(sample_at(comps[0], x, y));
                continue;
            }
            const float Y = static_cast<float>(sample_at(comps[0], x, y));
            const float Cb = static_cast<float>(sample_at(comps[1], x, y));
            const float Cr = static_cast<float>(sample_at(comps[2], x, y));
            float r, g, b;
            if (adobe_rgb) {
                r = Y;
                g = Cb;
                b = Cr;
            } else {  // JFIF YCbCr -> RGB
                r = Y + 1.402f * (Cr - 128.0f);
                g = Y - 0.344136f * (Cb - 128.0f) - 0.714136f * (Cr - 128.0f);
                b = Y + 1.772f * (Cb - 128.0f);
            }
            o[0] = static_cast<uint8_t>(std::clamp(static_cast<int>(std::lround(r)), 0, 255));
            o[1] = static_cast<uint8_t>(std::clamp(static_cast<int>(std::lround(g)), 0, 255));
            o[2] = static_cast<uint8_t>(std::clamp(static_cast<int>(std::lround(b)), 0, 255));
        }
    return true;
}

bool decode_image(const uint8_t* data, size_t n, Image8& out, std::string* err) {
    if (n >= 8 && data[0] == 137 && data[1] == 'P' && data[2] == 'N' && data[3] == 'G')
        return decode_png(data, n, out, err);
    if (n >= 3 && data[0] == 0xFF && data[1] == 0xD8) return decode_jpeg(data, n, out, err);
    if (err) *err = "unknown image format (PNG and JPEG are supported)";
    return false;
}

// ---------------------------------------------------------------- encoders --
namespace {

void put32be(std::vector<uint8_t>& o, uint32_t v) {
    o.push_back(static_cast<uint8_t>(v >> 24));
    o.push_back(static_cast<uint8_t>(v >> 16));
    o.push_back(static_cast<uint8_t>(v >> 8));
    o.push_back(static_cast<uint8_t>(v));
}
void put16be(std::vector<uint8_t>& o, uint32_t v) {
    o.push_back(static_cast<uint8_t>(v >> 8));
    o.push_back(static_cast<uint8_t>(v));
}
void put32le(std::vector<uint8_t>& o, uint32_t v) {
    for (int i = 0; i < 4; ++i) o.push_back(static_cast<uint8_t>(v >> (8 * i)));
}
void put16le(std::vector<uint8_t>& o, uint32_t v) {
    o.push_back(static_cast<uint8_t>(v));
    o.push_back(static_cast<uint8_t>(v >> 8));
}

void png_chunk(std::vector<uint8_t>& o, const char* type, const std::vector<uint8_t>& body) {
    put32be(o, static_cast<uint32_t>(body.size()));
    const size_t at = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), body.begin(), body.end());
    put32be(o, crc32(o.data() + at, body.size() + 4) ^ 0xFFFFFFFFu);
}

}  // namespace

// Rows filtered with "Up" (type 2), stored in deflate blocks without
// compression (RFC 1951 3.2.4): a valid PNG any reader decodes exactly.
void encode_png(int w, int h, const uint8_t* rgb, std::vector<uint8_t>& out) {
    out.clear();
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    out.insert(out.end(), sig, sig + 8);
    std::vector<uint8_t> ihdr;
    put32be(ihdr, static_cast<uint32_t>(w));
    put32be(ihdr, static_cast<uint32_t>(h));
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
    png_chunk(out, "IHDR", ihdr);
    const size_t stride = static_cast<size_t>(w) * 3;
    std::vector<uint8_t> raw;
    raw.reserve((stride + 1) * h);
    for (int y = 0; y < h; ++y) {
        raw.push_back(2);
        const uint8_t* row = rgb + stride * y;
        const uint8_t* up = y ? rgb + stride * (y - 1) : nullptr;
        for (size_t i = 0; i < stride; ++i) raw.push_back(static_cast<uint8_t>(row[i] - (up ? up[i] : 0)));
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    size_t off = 0;
    do {
        const size_t len = std::min<size_t>(65535, raw.size() - off);
        z.push_back(off + len == raw.size() ? 1 : 0);
        put16le(z, static_cast<uint32_t>(len));
        put16le(z, static_cast<uint32_t>(len ^ 0xFFFFu));
        z.insert(z.end(), raw.begin() + static_cast<std::ptrdiff_t>(off),
                 raw.begin() + static_cast<std::ptrdiff_t>(off + len));
        off += len;
    } while (off < raw.size());
    put32be(z, adler32(raw.data(), raw.size()));
    png_chunk(out, "IDAT", z);
    png_chunk(out, "IEND", {});
}

void encode_bmp(int w, int h, const uint8_t* rgb, std::vector<uint8_t>& out) {
    out.clear();
    const uint32_t stride = (static_cast<uint32_t>(w) * 3 + 3) & ~3u;
    const uint32_t size = 54 + stride * static_cast<uint32_t>(h);
    out.push_back('B');
    out.push_back('M');
    put32le(out, size);
    put32le(out, 0);
    put32le(out, 54);
    put32le(out, 40);
    put32le(out, static_cast<uint32_t>(w));
    put32le(out, static_cast<uint32_t>(h));  // positive: bottom row first
    put16le(out, 1);
    put16le(out, 24);
    put32le(out, 0);
    put32le(out, stride * static_cast<uint32_t>(h));
    put32le(out, 2835);
    put32le(out, 2835);
    put32le(out, 0);
    put32le(out, 0);
    for (int y = h - 1; y >= 0; --y) {
        const uint8_t* row = rgb + static_cast<size_t>(y) * w * 3;
        for (int x = 0; x < w; ++x) {
            out.push_back(row[x * 3 + 2]);
            out.push_back(row[x * 3 + 1]);
            out.push_back(row[x * 3 + 0]);
        }
        for (uint32_t p = static_cast<uint32_t>(w) * 3; p < stride; ++p) out.push_back(0);
    }
}

void encode_tga(int w, int h, const uint8_t* rgb, std::vector<uint8_t>& out) {
    out.clear();
    out.insert(out.end(), {0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0});  // uncompressed true colour
    put16le(out, static_cast<uint32_t>(w));
    put16le(out, static_cast<uint32_t>(h));
    out.push_back(24);
    out.push_back(0x20);  // first row is the top row
    for (size_t i = 0; i < static_cast<size_t>(w) * h; ++i) {
        out.push_back(rgb[i * 3 + 2]);
        out.push_back(rgb[i * 3 + 1]);
        out.push_back(rgb[i * 3 + 0]);
    }
}

namespace {

// Huffman code lengths limited to 16 bits for symbol frequencies (ITU T.81
// K.2: a reserved symbol keeps the all-ones code out of use).
void jpeg_code_lengths(const std::vector<long>& freq_in, uint8_t bits[17], std::vector<uint8_t>& vals) {
    const int N = static_cast<int>(freq_in.size());
    std::vector<long> freq(freq_in);
    freq.push_back(1);  // reserved
    std::vector<int> codesize(N + 1, 0), others(N + 1, -1);
    while (true) {
        int v1 = -1, v2 = -1;
        for (int i = 0; i <= N; ++i)
            if (freq[i] > 0 && (v1 < 0 || freq[i] <= freq[v1])) v1 = i;
        for (int i = 0; i <= N; ++i)
            if (freq[i] > 0 && i != v1 && (v2 < 0 || freq[i] <= freq[v2])) v2 = i;
        if (v2 < 0) break;
        freq[v1] += freq[v2];
        freq[v2] = 0;
        ++codesize[v1];
        while (others[v1] >= 0) {
            v1 = others[v1];
            ++codesize[v1];
        }
        others[v1] = v2;
        ++codesize[v2];
        while (others[v2] >= 0) {
            v2 = others[v2];
            ++codesize[v2];
        }
    }
    int cnt[33] = {};
    for (int i = 0; i <= N; ++i)
        if (codesize[i]) cnt[std::min(codesize[i], 32)]++;
    for (int i = 32; i > 16; --i)
        while (cnt[i] > 0) {
            int j = i - 2;
            while (cnt[j] == 0) --j;
            cnt[i] -= 2;
            cnt[i - 1] += 1;
            cnt[j + 1] += 2;
            cnt[j] -= 1;
        }
    int i = 16;
    while (cnt[i] == 0) --i;
    cnt[i] -= 1;  // drop the reserved symbol
    for (int l = 1; l <= 16; ++l) bits[l] = static_cast<uint8_t>(cnt[l]);
    vals.clear();
    for (int l = 1; l <= 32; ++l)
        for (int s = 0; s < N; ++s)
            if (codesize[s] == l) vals.push_back(static_cast<uint8_t>(s));
}

struct JCode {
    uint16_t code[256] = {};
    uint8_t len[256] = {};
    void build(const uint8_t bits[17], const std::vector<uint8_t>& vals) {
        int code_v = 0;
        size_t k = 0;
        for (int l = 1; l <= 16; ++l) {
            for (int i = 0; i < bits[l]; ++i, ++k, ++code_v) {
                code[vals[k]] = static_cast<uint16_t>(code_v);
                len[vals[k]] = static_cast<uint8_t>(l);
            }
            code_v <<= 1;
        }
    }
};

struct BitOut {
    std::vector<uint8_t>& o;
    uint32_t buf = 0;
    int cnt = 0;
    explicit BitOut(std::vector<uint8_t>& out) : o(out) {}
    void put(uint32_t v, int n) {
        for (int i = n - 1; i >= 0; --i) {
            buf = (buf << 1) | ((v >> i) & 1u);
            if (++cnt == 8) {
                o.push_back(static_cast<uint8_t>(buf));
                if (buf == 0xFF) o.push_back(0);
                buf = 0;
                cnt = 0;
            }
        }
    }
    void flush() {
        while (cnt) put(1, 1);
    }
};

int bit_size(int v) {
    v = std::abs(v);
    int s = 0;
    while (v) {
        ++s;
        v >>= 1;
    }
    return s;
}

}  // namespace

// Baseline JPEG, 4:4:4 YCbCr, the example quantization tables of ITU T.81
// Annex K scaled by `quality` (1..100, the common quality mapping), Huffman
// tables fitted to the image.
void encode_jpeg(int w, int h, const uint8_t* rgb, int quality, std::vector<uint8_t>& out) {
    out.clear();
    static const uint8_t qbase[2][64] = {
        {16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55, 14, 13, 16, 24, 40, 57, 69, 56,
         14, 17, 22, 29, 51, 87, 80, 62, 18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64, 81, 104, 113, 92,
         49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99},
        {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99, 24, 26, 56, 99, 99, 99, 99, 99,
         47, 66, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
         99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99}};
    quality = std::clamp(quality, 1, 100);
    const int scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;
    const auto& zz = zigzag();
    uint8_t q[2][64];  // natural order
    for (int t = 0; t < 2; ++t)
        for (int i = 0; i < 64; ++i) q[t][i] = static_cast<uint8_t>(std::clamp((qbase[t][i] * scale + 50) / 100, 1, 255));
    // forward DCT of every block (T.81 A.3.3), quantized, zigzag order
    const int bw = (w + 7) / 8, bh = (h + 7) / 8;
    std::vector<int16_t> coef(static_cast<size_t>(bw) * bh * 3 * 64);
    static const std::array<double, 64> cosines = [] {
        std::array<double, 64> t{};
        for (int x = 0; x < 8; ++x)
            for (int u = 0; u < 8; ++u)
                t[x * 8 + u] = (u == 0 ? std::sqrt(0.5) : 1.0) * std::cos((2 * x + 1) * u * 3.14159265358979323846 / 16.0) * 0.5;
        return t;
    }();
    for (int by = 0; by < bh; ++by)
        for (int bx = 0; bx < bw; ++bx) {
            double ycc[3][64];
            for (int y = 0; y < 8; ++y)
                for (int x = 0; x < 8; ++x) {
                    const int sx = std::min(bx * 8 + x, w - 1), sy = std::min(by * 8 + y, h - 1);
                    const uint8_t* p = rgb + (static_cast<size_t>(sy) * w + sx) * 3;
                    const double r = p[0], g = p[1], b = p[2];
                    ycc[0][y * 8 + x] = 0.299 * r + 0.587 * g + 0.114 * b - 128.0;
                    ycc[1][y * 8 + x] = -0.168736 * r - 0.331264 * g + 0.5 * b;
                    ycc[2][y * 8 + x] = 0.5 * r - 0.418688 * g - 0.081312 * b;
                }
            for (int c = 0; c < 3; ++c) {
                double tmp[64];
                for (int y = 0; y < 8; ++y)
                    for (int u = 0; u < 8; ++u) {
                        double s = 0.0;
                        for (int x = 0; x < 8; ++x) s += cosines[x * 8 + u] * ycc[c][y * 8 + x];
                        tmp[y * 8 + u] = s;
                    }
                int16_t* blk = coef.data() + ((static_cast<size_t>(by) * bw + bx) * 3 + c) * 64;
                for (int v = 0; v < 8; ++v)
                    for (int u = 0; u < 8; ++u) {
                        double s = 0.0;
                        for (int y = 0; y < 8; ++y) s += cosines[y * 8 + v] * tmp[y * 8 + u];
                        const int nat = v * 8 + u;
                        blk[nat] = static_cast<int16_t>(std::lround(s / q[c ? 1 : 0][nat]));
                    }
            }
        }
    // symbol statistics -> tables
    std::vector<long> fdc[2] = {std::vector<long>(12, 0), std::vector<long>(12, 0)};
    std::vector<long> fac[2] = {std::vector<long>(256, 0), std::vector<long>(256, 0)};
    auto walk = [&](auto&& emit_dc, auto&& emit_ac) {
        int pred[3] = {0, 0, 0};
        for (size_t b = 0; b < static_cast<size_t>(bw) * bh; ++b)
            for (int c = 0; c < 3; ++c) {
                const int16_t* blk = coef.data() + (b * 3 + c) * 64;
                const int t = c ? 1 : 0;
                const int diff = blk[0] - pred[c];
                pred[c] = blk[0];
                emit_dc(t, diff);
                int run = 0;
                for (int k = 1; k < 64; ++k) {
                    const int v = blk[zz[k]];
                    if (v == 0) {
                        ++run;
                        continue;
                    }
                    while (run > 15) {
                        emit_ac(t, 0xF0, 0);
                        run -= 16;
                    }
                    emit_ac(t, (run << 4) | bit_size(v), v);
                    run = 0;
                }
                if (run) emit_ac(t, 0x00, 0);
            }
    };
    walk([&](int t, int diff) { fdc[t][bit_size(diff)]++; }, [&](int t, int rs, int) { fac[t][rs]++; });
    uint8_t dcbits[2][17], acbits[2][17];
    std::vector<uint8_t> dcvals[2], acvals[2];
    JCode dcc[2], acc[2];
    for (int t = 0; t < 2; ++t) {
        jpeg_code_lengths(fdc[t], dcbits[t], dcvals[t]);
        jpeg_code_lengths(fac[t], acbits[t], acvals[t]);
        dcc[t].build(dcbits[t], dcvals[t]);
        acc[t].build(acbits[t], acvals[t]);
    }
    // headers
    out.insert(out.end(), {0xFF, 0xD8, 0xFF, 0xE0, 0, 16, 'J', 'F', 'I', 'F', 0, 1, 1, 0, 0, 1, 0, 1, 0, 0});
    for (int t = 0; t < 2; ++t) {
        out.insert(out.end(), {0xFF, 0xDB, 0, 67, static_cast<uint8_t>(t)});
        for (int k = 0; k < 64; ++k) out.push_back(q[t][zz[k]]);
    }
    out.insert(out.end(), {0xFF, 0xC0, 0, 17, 8});
    put16be(out, static_cast<uint32_t>(h));
    put16be(out, static_cast<uint32_t>(w));
    out.insert(out.end(), {3, 1, 0x11, 0, 2, 0x11, 1, 3, 0x11, 1});
    auto dht = [&](int cls, int id, const uint8_t bits[17], const std::vector<uint8_t>& vals) {
        out.insert(out.end(), {0xFF, 0xC4});
        put16be(out, static_cast<uint32_t>(3 + 16 + vals.size()));
        out.push_back(static_cast<uint8_t>((cls << 4) | id));
        for (int l = 1; l <= 16; ++l) out.push_back(bits[l]);
        out.insert(out.end(), vals.begin(), vals.end());
    };
    for (int t = 0; t < 2; ++t) {
        dht(0, t, dcbits[t], dcvals[t]);
        dht(1, t, acbits[t], acvals[t]);
    }
    out.insert(out.end(), {0xFF, 0xDA, 0, 12, 3, 1, 0x00, 2, 0x11, 3, 0x11, 0, 63, 0});
    BitOut bo(out);
    walk(
        [&](int t, int diff) {
            const int s = bit_size(diff);
            bo.put(dcc[t].code[s], dcc[t].len[s]);
            if (s) bo.put(static_cast<uint32_t>(diff < 0 ? diff - 1 : diff) & ((1u << s) - 1u), s);
        },
        [&](int t, int rs, int v) {
            bo.put(acc[t].code[rs], acc[t].len[rs]);
            const int s = rs & 15;
            if (s) bo.put(static_cast<uint32_t>(v < 0 ? v - 1 : v) & ((1u << s) - 1u), s);
        });
    bo.flush();
    out.insert(out.end(), {0xFF, 0xD9});
}

}  // namespace ptamd
