// image.cpp -- PNG (zlib inflate + the PNG filter and interlace rules of the
// PNG specification, ISO/IEC 15948) and BMP decoders of the asset loader.
#include "image.h"

#include <zlib.h>

#include <cstring>

namespace sw_img {

namespace {

uint32_t be32(const uint8_t *p)
{
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}

inline uint32_t bgra(uint32_t r, uint32_t g, uint32_t b, uint32_t a)
{
    return b | g << 8 | r << 16 | a << 24;
}

int paeth(int a, int b, int c)
{
    int p = a + b - c, pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
    if (pa <= pb && pa <= pc)
        return a;
    return pb <= pc ? b : c;
}

// Undo the per-row filters of one (sub)image in place; rows are
// [filter byte][stride bytes]; returns false for an unknown filter type.
bool unfilter(uint8_t *rows, size_t height, size_t stride, size_t bpp, std::vector<uint8_t> &img)
{
    img.assign(height * stride, 0);
    for (size_t y = 0; y < height; y++) {
        const uint8_t *src = rows + y * (stride + 1);
        const int f = src[0];
        src++;
        uint8_t *cur = &img[y * stride];
        const uint8_t *prev = y ? &img[(y - 1) * stride] : nullptr;
        for (size_t i = 0; i < stride; i++) {
            const int a = i >= bpp ? cur[i - bpp] : 0;
            const int b = prev ? prev[i] : 0;
            const int c = (prev && i >= bpp) ? prev[i - bpp] : 0;
            int v = src[i];
            switch (f) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) >> 1; break;
            case 4: v += paeth(a, b, c); break;
            default: return false;
            }
            cur[i] = (uint8_t)v;
        }
    }
    return true;
}

struct Png {
    int w = 0, h = 0, depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> plte;   // rgb triples
    std::vector<uint8_t> trns;   // raw tRNS chunk
    int channels() const { return ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4; }

    // sample k of pixel x in an unfiltered row, at the file's bit depth
    uint32_t sample(const uint8_t *row, int x, int k) const
    {
        const int ch = channels();
        if (depth == 8)
            return row[x * ch + k];
        if (depth == 16)
            return (uint32_t)row[2 * (x * ch + k)] << 8 | row[2 * (x * ch + k) + 1];
        const int bit = x * depth;   // one channel when depth < 8
        return (uint32_t)(row[bit >> 3] >> (8 - depth - (bit & 7))) & ((1u << depth) - 1);
    }
    // top byte of a 16-bit sample (png_set_strip_16); gray below 8 bits
    // scaled by 255 / (2^depth - 1) (png_set_expand_gray_1_2_4_to_8)
    uint32_t to8(uint32_t v) const
    {
        switch (depth) {
        case 1: return v * 0xFF;
        case 2: return v * 0x55;
        case 4: return v * 0x11;
        case 16: return v >> 8;
        default: return v;
        }
    }
    uint32_t pixel(const uint8_t *row, int x) const
    {
        switch (ctype) {
        case 0: {   // gray (+ tRNS gray key compared at the file's depth)
            const uint32_t g = sample(row, x, 0);
            uint32_t a = 255;
            if (trns.size() >= 2 && g == ((uint32_t)trns[0] << 8 | trns[1]))
                a = 0;
            const uint32_t g8 = to8(g);
            return bgra(g8, g8, g8, a);
        }
        case 2: {   // rgb (+ tRNS rgb key)
            const uint32_t r = sample(row, x, 0), g = sample(row, x, 1), b = sample(row, x, 2);
            uint32_t a = 255;
            if (trns.size() >= 6 && r == ((uint32_t)trns[0] << 8 | trns[1]) &&
                g == ((uint32_t)trns[2] << 8 | trns[3]) && b == ((uint32_t)trns[4] << 8 | trns[5]))
                a = 0;
            return bgra(to8(r), to8(g), to8(b), a);
        }
        case 3: {   // palette: entries past PLTE read as black, alpha from tRNS
            const uint32_t i = sample(row, x, 0);
            uint32_t r = 0, g = 0, b = 0, a = 255;
            if (3 * i + 2 < plte.size()) {
                r = plte[3 * i];
                g = plte[3 * i + 1];
                b = plte[3 * i + 2];
            }
            if (i < trns.size())
                a = trns[i];
            return bgra(r, g, b, a);
        }
        case 4: {   // gray + alpha
            const uint32_t g8 = to8(sample(row, x, 0));
            return bgra(g8, g8, g8, to8(sample(row, x, 1)));
        }
        default:    // rgba
            return bgra(to8(sample(row, x, 0)), to8(sample(row, x, 1)), to8(sample(row, x, 2)),
                        to8(sample(row, x, 3)));
        }
    }
};

}  // namespace

bool decode_png(const uint8_t *data, size_t n, std::vector<uint32_t> &out, int &w, int &h, std::string &err)
{
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    if (n < 8 || memcmp(data, sig, 8) != 0) {
        err = "not a PNG file";
        return false;
    }
    Png P;
    std::vector<uint8_t> z;
    bool have_ihdr = false, have_iend = false;
    size_t pos = 8;
    while (pos + 12 <= n && !have_iend) {
        const uint32_t len = be32(data + pos);
        if (len > n - pos - 12) {
            err = "PNG chunk runs past the end of the data";
            return false;
        }
        const uint8_t *type = data + pos + 4, *body = data + pos + 8;
        const bool critical = !(type[0] & 0x20);
        const uint32_t crc = (uint32_t)crc32(crc32(0L, Z_NULL, 0), type, len + 4);
        pos += 12 + (size_t)len;
        if (crc != be32(body + len)) {
            if (critical) {
                err = "PNG chunk CRC mismatch";
                return false;
            }
            continue;   // libpng's default: a damaged ancillary chunk is dropped
        }
        if (!memcmp(type, "IHDR", 4)) {
            if (len != 13) {
                err = "bad IHDR";
                return false;
            }
            P.w = (int)be32(body);
            P.h = (int)be32(body + 4);
            P.depth = body[8];
            P.ctype = body[9];
            P.interlace = body[12];
            have_ihdr = true;
            if (body[10] != 0 || body[11] != 0 || P.interlace > 1) {
                err = "unsupported PNG compression, filter or interlace method";
                return false;
            }
            const int d = P.depth, c = P.ctype;
            const bool ok = (c == 0 && (d == 1 || d == 2 || d == 4 || d == 8 || d == 16)) ||
                            (c == 3 && (d == 1 || d == 2 || d == 4 || d == 8)) ||
                            ((c == 2 || c == 4 || c == 6) && (d == 8 || d == 16));
            if (!ok || P.w <= 0 || P.h <= 0 || (int64_t)P.w * P.h > (1ll << 28)) {
                err = "invalid PNG header (size, colour type or bit depth)";
                return false;
            }
        } else if (!memcmp(type, "PLTE", 4)) {
            P.plte.assign(body, body + len);
        } else if (!memcmp(type, "tRNS", 4)) {
            P.trns.assign(body, body + len);
        } else if (!memcmp(type, "IDAT", 4)) {
            z.insert(z.end(), body, body + len);
        } else if (!memcmp(type, "IEND", 4)) {
            have_iend = true;
        } else if (critical) {
            err = "unknown critical PNG chunk";
            return false;
        }
    }
    if (!have_ihdr || z.empty()) {
        err = "PNG without IHDR or image data";
        return false;
    }
    if (P.ctype == 3 && P.plte.empty()) {
        err = "palette PNG without PLTE";
        return false;
    }
    // the 7 Adam7 passes (x0, y0, dx, dy), or one pass for the whole image
    static const int adam7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                    {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const int one[1][4] = {{0, 0, 1, 1}};
    const int (*passes)[4] = P.interlace ? adam7 : one;
    const int npass = P.interlace ? 7 : 1;
    const size_t bits = (size_t)P.channels() * P.depth, bpp = bits < 8 ? 1 : bits / 8;
    size_t raw_size = 0;
    for (int k = 0; k < npass; k++) {
        const size_t pw = ((size_t)P.w - passes[k][0] + passes[k][2] - 1) / passes[k][2];
        const size_t ph = ((size_t)P.h - passes[k][1] + passes[k][3] - 1) / passes[k][3];
        if (pw && ph && P.w > passes[k][0] && P.h > passes[k][1])
            raw_size += ph * (1 + (pw * bits + 7) / 8);
    }
    std::vector<uint8_t> raw(raw_size);
    z_stream zs;
    memset(&zs, 0, sizeof zs);
    if (inflateInit(&zs) != Z_OK) {
        err = "zlib init failed";
        return false;
    }
    zs.next_in = z.data();
    zs.avail_in = (uInt)z.size();
    zs.next_out = raw.data();
    zs.avail_out = (uInt)raw.size();
    const int zr = inflate(&zs, Z_FINISH);
    const size_t got = raw.size() - zs.avail_out;
    inflateEnd(&zs);
    if ((zr != Z_STREAM_END && zr != Z_BUF_ERROR && zr != Z_OK) || got != raw.size()) {
        err = "PNG image data is truncated or corrupt";
        return false;
    }
    w = P.w;
    h = P.h;
    out.assign((size_t)P.w * P.h, 0);
    size_t off = 0;
    std::vector<uint8_t> img;
    for (int k = 0; k < npass; k++) {
        const int x0 = passes[k][0], y0 = passes[k][1], dx = passes[k][2], dy = passes[k][3];
        if (P.w <= x0 || P.h <= y0)
            continue;
        const size_t pw = ((size_t)P.w - x0 + dx - 1) / dx, ph = ((size_t)P.h - y0 + dy - 1) / dy;
        const size_t stride = (pw * bits + 7) / 8;
        if (!unfilter(&raw[off], ph, stride, bpp, img)) {
            err = "unknown PNG filter type";
            return false;
        }
        off += ph * (stride + 1);
        for (size_t y = 0; y < ph; y++)
            for (size_t x = 0; x < pw; x++)
                out[(y0 + y * dy) * (size_t)P.w + x0 + x * dx] = P.pixel(&img[y * stride], (int)x);
    }
    return true;
}

bool decode_bmp(const uint8_t *data, size_t n, std::vector<uint32_t> &out, int &w, int &h, std::string &err)
{
    if (n < 54) {
        err = "BMP shorter than its 54-byte header";
        return false;
    }
    int32_t width, height;
    memcpy(&width, data + 18, 4);
    memcpy(&height, data + 22, 4);
    if (width <= 0 || height <= 0 || (int64_t)width * height > (1ll << 28)) {
        err = "BMP size not supported";
        return false;
    }
    const size_t pad = (size_t)(width * 3) % 4;   // image.cpp:78, as read (not the 4-byte alignment)
    const size_t need = 54 + (size_t)height * ((size_t)width * 3 + pad) - pad;
    if (n < need) {
        err = "BMP pixel data is truncated";
        return false;
    }
    w = width;
    h = height;
    out.assign((size_t)width * height, 0);
    const uint8_t *p = data + 54;
    for (int row = height - 1; row >= 0; row--) {   // the last row of the image comes first
        for (int i = 0; i < width; i++, p += 3)
            out[(size_t)row * width + i] = bgra(p[2], p[1], p[0], 255);
        if (row)
            p += pad;
    }
    return true;
}

}  // namespace sw_img
