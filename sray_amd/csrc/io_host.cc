/*
 * io_host.cc — host-side input/output of the path: the photon-map text
 * dump and restore (PhotonMap::dump / restore, src/pmap.cc:180-251), the
 * framebuffer written as PFM or PNG (Image::save, src/img.cc:80-83), and the
 * 32-byte status records of the reference's backend protocol (emit_status,
 * src/render.cc:279-289), which the serving engine also writes.
 */
#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include <zlib.h>

#include "host.h"

namespace {
thread_local std::string g_io_err;

int io_fail(int code, const std::string &msg)
{
    g_io_err = msg;
    return code;
}
} // namespace

namespace sray {
namespace {

bool read_file(const std::string &path, std::vector<unsigned char> &buf, std::string &err)
{
    FILE *fp = fopen(path.c_str(), "rb");
    if (!fp) {
        err = "failed to open image: " + path + ": " + strerror(errno);
        return false;
    }
    unsigned char tmp[65536];
    size_t n;
    try {
        while ((n = fread(tmp, 1, sizeof tmp, fp)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    } catch (const std::bad_alloc &) {
        fclose(fp);
        err = "image too large for host memory: " + path;
        return false;
    }
    bool bad = ferror(fp) != 0;
    fclose(fp);
    if (bad) err = "failed to read image: " + path;
    return !bad;
}

/* netpbm header tokens: whitespace-separated, '#' starts a comment to the
 * end of the line */
bool pnm_token(const std::vector<unsigned char> &b, size_t &at, std::string &tok)
{
    tok.clear();
    while (at < b.size()) {
        if (b[at] == '#') {
            while (at < b.size() && b[at] != '\n') at++;
        } else if (isspace(b[at])) at++;
        else break;
    }
    while (at < b.size() && !isspace(b[at]) && b[at] != '#') tok.push_back((char)b[at++]);
    return !tok.empty();
}
bool pnm_int(const std::vector<unsigned char> &b, size_t &at, long &v)
{
    std::string t;
    if (!pnm_token(b, at, t)) return false;
    char *end = nullptr;
    v = strtol(t.c_str(), &end, 10);
    return *end == 0;
}

constexpr long MAX_SIDE = 1L << 15; /* 32768 texels a side */

/* P2/P3 (text) and P5/P6 (binary) grey / RGB, maxval up to 65535 (two
 * bytes, most significant first, above 255): value / maxval, alpha 1 */
bool read_pnm(const std::vector<unsigned char> &b, const std::string &path, int &w, int &h, std::vector<float> &px,
              std::string &err)
{
    const char kind = (char)b[1];
    const int ch = (kind == '3' || kind == '6') ? 3 : 1;
    const bool text = kind == '2' || kind == '3';
    size_t at = 2;
    long W, H, maxv;
    if (!pnm_int(b, at, W) || !pnm_int(b, at, H) || !pnm_int(b, at, maxv) || W <= 0 || H <= 0 || W > MAX_SIDE ||
        H > MAX_SIDE || maxv <= 0 || maxv > 65535) {
        err = "bad netpbm header: " + path;
        return false;
    }
    at++; /* the single whitespace byte after maxval */
    const size_t n = (size_t)W * (size_t)H * ch, bps = maxv > 255 ? 2 : 1;
    if (!text && b.size() < at + n * bps) {
        err = "truncated netpbm image: " + path;
        return false;
    }
    px.assign((size_t)W * H * 4, 1.0f);
    const float scale = (float)maxv;
    for (size_t i = 0; i < n; i++) {
        long v;
        if (text) {
            if (!pnm_int(b, at, v)) {
                err = "truncated netpbm image: " + path;
                return false;
            }
        } else v = bps == 2 ? ((long)b[at + 2 * i] << 8) | b[at + 2 * i + 1] : (long)b[at + i];
        const float f = (float)v / scale;
        const size_t p = i / ch;
        if (ch == 3) px[p * 4 + i % 3] = f;
        else px[p * 4] = px[p * 4 + 1] = px[p * 4 + 2] = f;
    }
    w = (int)W;
    h = (int)H;
    return true;
}

/* PFM: "PF" (RGB) or "Pf" (grey), width height, a scale whose sign gives
 * the byte order (negative: little-endian), float32 rows bottom row first */
bool read_pfm(const std::vector<unsigned char> &b, const std::string &path, int &w, int &h, std::vector<float> &px,
              std::string &err)
{
    const int ch = b[1] == 'F' ? 3 : 1;
    size_t at = 2;
    long W, H;
    std::string st;
    if (!pnm_int(b, at, W) || !pnm_int(b, at, H) || !pnm_token(b, at, st) || W <= 0 || H <= 0 || W > MAX_SIDE ||
        H > MAX_SIDE) {
        err = "bad PFM header: " + path;
        return false;
    }
    const double scale = strtod(st.c_str(), nullptr);
    if (scale == 0.0) {
        err = "bad PFM header: " + path;
        return false;
    }
    at++;
    const bool little = scale < 0.0;
    const size_t n = (size_t)W * (size_t)H * ch;
    if (b.size() < at + n * 4) {
        err = "truncated PFM image: " + path;
        return false;
    }
    px.assign((size_t)W * H * 4, 1.0f);
    for (size_t i = 0; i < n; i++) {
        const unsigned char *q = &b[at + 4 * i];
        uint32_t u = little ? (uint32_t)q[0] | (uint32_t)q[1] << 8 | (uint32_t)q[2] << 16 | (uint32_t)q[3] << 24
                            : (uint32_t)q[3] | (uint32_t)q[2] << 8 | (uint32_t)q[1] << 16 | (uint32_t)q[0] << 24;
        float f;
        memcpy(&f, &u, 4);
        const size_t pf = i / ch, x = pf % (size_t)W, yb = pf / (size_t)W;
        const size_t p = ((size_t)H - 1 - yb) * (size_t)W + x; /* top row first */
        if (ch == 3) px[p * 4 + i % 3] = f;
        else px[p * 4] = px[p * 4 + 1] = px[p * 4 + 2] = f;
    }
    w = (int)W;
    h = (int)H;
    return true;
}

uint32_t be32(const unsigned char *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

/* PNG (ISO/IEC 15948): grey, RGB, palette, grey + alpha and RGBA at every
 * bit depth, tRNS transparency, filters 0-4; the IDAT stream inflated with
 * zlib.  Interlaced (Adam7) files are refused.  A channel of depth d reads
 * value / (2^d - 1). */
bool read_png(const std::vector<unsigned char> &b, const std::string &path, int &w, int &h, std::vector<float> &px,
              std::string &err)
{
    size_t at = 8;
    uint32_t W = 0, H = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<unsigned char> idat, plte, trns;
    bool end = false;
    while (!end) {
        if (at + 12 > b.size()) {
            err = "truncated PNG: " + path;
            return false;
        }
        const uint32_t len = be32(&b[at]);
        const std::string type((const char *)&b[at + 4], 4);
        if (len > b.size() - at - 12) {
            err = "truncated PNG: " + path;
            return false;
        }
        const unsigned char *d = &b[at + 8];
        uLong crc = crc32(0L, &b[at + 4], 4 + len);
        if ((uint32_t)crc != be32(d + len)) {
            err = "PNG chunk " + type + " fails its CRC: " + path;
            return false;
        }
        if (type == "IHDR") {
            if (len < 13) {
                err = "bad PNG header: " + path;
                return false;
            }
            W = be32(d);
            H = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
            if (d[10] != 0 || d[11] != 0) {
                err = "unknown PNG compression or filter method: " + path;
                return false;
            }
        } else if (type == "PLTE") plte.assign(d, d + len);
        else if (type == "tRNS") trns.assign(d, d + len);
        else if (type == "IDAT") idat.insert(idat.end(), d, d + len);
        else if (type == "IEND") end = true;
        at += 12 + (size_t)len;
    }
    static const int chans[7] = {1, 0, 3, 1, 2, 0, 4};
    const bool okdepth = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                         (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                         ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (W == 0 || H == 0 || W > (uint32_t)MAX_SIDE || H > (uint32_t)MAX_SIDE || !okdepth) {
        err = "unsupported PNG size, colour type or depth: " + path;
        return false;
    }
    if (interlace != 0) {
        err = "interlaced PNG not supported: " + path;
        return false;
    }
    if (ctype == 3 && (plte.empty() || plte.size() % 3 != 0)) {
        err = "PNG palette missing: " + path;
        return false;
    }
    const int nch = chans[ctype];
    const size_t bpp = ((size_t)nch * depth + 7) / 8; /* filter unit, at least one byte */
    const size_t stride = ((size_t)W * nch * depth + 7) / 8;
    std::vector<unsigned char> raw;
    try {
        raw.resize((stride + 1) * H);
    } catch (const std::bad_alloc &) {
        err = "image too large for host memory: " + path;
        return false;
    }
    uLongf rlen = (uLongf)raw.size();
    int zr = uncompress(raw.data(), &rlen, idat.data(), (uLong)idat.size());
    if (zr != Z_OK || rlen != raw.size()) {
        err = "corrupt PNG image data: " + path;
        return false;
    }
    /* undo the scanline filters in place */
    std::vector<unsigned char> prev(stride, 0);
    for (uint32_t y = 0; y < H; y++) {
        unsigned char *row = &raw[(size_t)y * (stride + 1)];
        const int f = row[0];
        unsigned char *c = row + 1;
        for (size_t i = 0; i < stride; i++) {
            const int a = i >= bpp ? c[i - bpp] : 0, up = prev[i], ul = i >= bpp ? prev[i - bpp] : 0;
            int v = c[i];
            switch (f) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += up; break;
            case 3: v += (a + up) >> 1; break;
            case 4: {
                const int p = a + up - ul, pa = abs(p - a), pb = abs(p - up), pc = abs(p - ul);
                v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? up : ul);
                break;
            }
            default:
                err = "bad PNG filter type: " + path;
                return false;
            }
            c[i] = (unsigned char)v;
        }
        memcpy(prev.data(), c, stride);
    }
    px.assign((size_t)W * H * 4, 1.0f);
    const float maxv = (float)((1u << depth) - 1u);
    auto sample = [&](const unsigned char *row, size_t k) -> uint32_t {
        if (depth == 16) return (uint32_t)row[2 * k] << 8 | row[2 * k + 1];
        if (depth == 8) return row[k];
        const size_t bit = k * (size_t)depth;
        return (uint32_t)(row[bit >> 3] >> (8 - depth - (int)(bit & 7))) & ((1u << depth) - 1u);
    };
    for (uint32_t y = 0; y < H; y++) {
        const unsigned char *row = &raw[(size_t)y * (stride + 1) + 1];
        for (uint32_t x = 0; x < W; x++) {
            float *o = &px[((size_t)y * W + x) * 4];
            if (ctype == 3) {
                const uint32_t ix = sample(row, x);
                if ((size_t)ix * 3 + 2 >= plte.size()) {
                    err = "PNG palette index out of range: " + path;
                    return false;
                }
                o[0] = (float)plte[ix * 3] / 255.0f;
                o[1] = (float)plte[ix * 3 + 1] / 255.0f;
                o[2] = (float)plte[ix * 3 + 2] / 255.0f;
                o[3] = ix < trns.size() ? (float)trns[ix] / 255.0f : 1.0f;
                continue;
            }
            uint32_t s[4];
            for (int k = 0; k < nch; k++) s[k] = sample(row, (size_t)x * nch + k);
            if (ctype == 0 || ctype == 4) {
                o[0] = o[1] = o[2] = (float)s[0] / maxv;
                if (ctype == 4) o[3] = (float)s[1] / maxv;
                else if (trns.size() >= 2 && s[0] == ((uint32_t)trns[0] << 8 | trns[1])) o[3] = 0.0f;
            } else {
                o[0] = (float)s[0] / maxv;
                o[1] = (float)s[1] / maxv;
                o[2] = (float)s[2] / maxv;
                if (ctype == 6) o[3] = (float)s[3] / maxv;
                else if (trns.size() >= 6 && s[0] == ((uint32_t)trns[0] << 8 | trns[1]) &&
                         s[1] == ((uint32_t)trns[2] << 8 | trns[3]) && s[2] == ((uint32_t)trns[4] << 8 | trns[5]))
                    o[3] = 0.0f;
            }
        }
    }
    w = (int)W;
    h = (int)H;
    return true;
}

} // namespace

bool read_image_rgbaf(const std::string &path, int &w, int &h, std::vector<float> &px, std::string &err)
{
    std::vector<unsigned char> b;
    if (!read_file(path, b, err)) return false;
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    try {
        if (b.size() >= 8 && memcmp(b.data(), sig, 8) == 0) return read_png(b, path, w, h, px, err);
        if (b.size() >= 3 && b[0] == 'P' && (b[1] == 'F' || b[1] == 'f')) return read_pfm(b, path, w, h, px, err);
        if (b.size() >= 3 && b[0] == 'P' && (b[1] == '2' || b[1] == '3' || b[1] == '5' || b[1] == '6'))
            return read_pnm(b, path, w, h, px, err);
    } catch (const std::bad_alloc &) {
        err = "image too large for host memory: " + path;
        return false;
    }
    err = "unknown image format (PNG, PPM/PGM and PFM are read): " + path;
    return false;
}

} // namespace sray

extern "C" {

const char *sray_io_last_error(void) { return g_io_err.c_str(); }

/* Image::load, src/img.cc:64-78 */
int sray_image_read(const char *path, float **rgba, int *width, int *height)
{
    if (!path || !rgba || !width || !height) return io_fail(SRAY_ERR_ARG, "sray_image_read: bad arguments");
    *rgba = nullptr;
    *width = *height = 0;
    std::vector<float> px;
    int w, h;
    std::string err;
    if (!sray::read_image_rgbaf(path, w, h, px, err)) return io_fail(SRAY_ERR_PARSE, err);
    float *buf = new (std::nothrow) float[px.size()];
    if (!buf) return io_fail(SRAY_ERR_NOMEM, "image too large for host memory");
    memcpy(buf, px.data(), px.size() * sizeof(float));
    *rgba = buf;
    *width = w;
    *height = h;
    return SRAY_OK;
}

void sray_image_free(float *rgba) { delete[] rgba; }

/* PhotonMap::dump, src/pmap.cc:180-205: the energy line, then one line per
 * photon with position, direction, normal and colour (rgb), printf "%f" */
int sray_photon_dump(const char *path, const sray_photon *p, uint32_t n, float energy)
{
    if (!path || (n && !p)) return io_fail(SRAY_ERR_ARG, "sray_photon_dump: bad arguments");
    FILE *fp = fopen(path, "w");
    if (!fp) return io_fail(SRAY_ERR_ARG, std::string("failed to dump photon map: ") + path + ": " + strerror(errno));
    bool ok = fprintf(fp, "energy %f\n", (double)energy) > 0;
    for (uint32_t i = 0; i < n && ok; i++) {
        const sray_photon &q = p[i];
        ok = fprintf(fp, "<%f %f %f> ", (double)q.pos[0], (double)q.pos[1], (double)q.pos[2]) > 0 &&
             fprintf(fp, "<%f %f %f> ", (double)q.dir[0], (double)q.dir[1], (double)q.dir[2]) > 0 &&
             fprintf(fp, "<%f %f %f> ", (double)q.norm[0], (double)q.norm[1], (double)q.norm[2]) > 0 &&
             fprintf(fp, "<%f %f %f>\n", (double)q.col[0], (double)q.col[1], (double)q.col[2]) > 0;
    }
    if (fclose(fp) != 0) ok = false;
    return ok ? SRAY_OK : io_fail(SRAY_ERR_ARG, std::string("failed to write photon dump: ") + path);
}

/* PhotonMap::restore, src/pmap.cc:207-251: reads the energy line and the
 * photons; every colour is scaled by energy / (the dump's energy), the
 * quotient taken in float and the product in double as the reference's
 * Vector3 * float, then stored as float */
int sray_photon_restore(const char *path, float energy, sray_photon **out, uint32_t *n)
{
    if (!path || !out || !n) return io_fail(SRAY_ERR_ARG, "sray_photon_restore: bad arguments");
    *out = nullptr;
    *n = 0;
    FILE *fp = fopen(path, "r");
    if (!fp) return io_fail(SRAY_ERR_ARG, std::string("failed to open photon dump: ") + path);
    float prev_energy = 0.0f;
    if (fscanf(fp, "energy %f\n", &prev_energy) < 1) {
        fclose(fp);
        return io_fail(SRAY_ERR_PARSE, std::string("failed to restore photon dump: ") + path);
    }
    const float scale = energy / prev_energy;
    std::vector<sray_photon> v;
    try {
        for (;;) {
            float x, y, z, dx, dy, dz, nx, ny, nz, cx, cy, cz;
            int res = fscanf(fp, "<%f %f %f> <%f %f %f> <%f %f %f> <%f %f %f>\n", &x, &y, &z, &dx, &dy, &dz, &nx,
                             &ny, &nz, &cx, &cy, &cz);
            if (res == EOF) break;
            if (res < 12) {
                fclose(fp);
                return io_fail(SRAY_ERR_PARSE, std::string("failed to restore photon dump: ") + path);
            }
            sray_photon q;
            q.pos[0] = x, q.pos[1] = y, q.pos[2] = z;
            q.dir[0] = dx, q.dir[1] = dy, q.dir[2] = dz;
            q.norm[0] = nx, q.norm[1] = ny, q.norm[2] = nz;
            q.col[0] = (float)((double)cx * (double)scale);
            q.col[1] = (float)((double)cy * (double)scale);
            q.col[2] = (float)((double)cz * (double)scale);
            v.push_back(q);
        }
    } catch (const std::bad_alloc &) {
        fclose(fp);
        return io_fail(SRAY_ERR_NOMEM, "photon dump too large for host memory");
    }
    fclose(fp);
    if (!v.empty()) {
        sray_photon *buf = new (std::nothrow) sray_photon[v.size()];
        if (!buf) return io_fail(SRAY_ERR_NOMEM, "photon dump too large for host memory");
        memcpy(buf, v.data(), v.size() * sizeof(sray_photon));
        *out = buf;
    }
    *n = (uint32_t)v.size();
    return SRAY_OK;
}

void sray_photon_free(sray_photon *p) { delete[] p; }

/* The float framebuffer (RGBA, row-major, top row first) as a PFM file:
 * "PF\n<w> <h>\n-1.0\n", then RGB float32 little-endian, bottom row first */
int sray_image_write_pfm(const char *path, const float *rgba, int width, int height)
{
    if (!path || !rgba || width <= 0 || height <= 0) return io_fail(SRAY_ERR_ARG, "sray_image_write_pfm: bad arguments");
    FILE *fp = fopen(path, "wb");
    if (!fp) return io_fail(SRAY_ERR_ARG, std::string("failed to open image file: ") + path + ": " + strerror(errno));
    bool ok = fprintf(fp, "PF\n%d %d\n-1.0\n", width, height) > 0;
    std::vector<float> row((size_t)width * 3);
    for (int y = height - 1; y >= 0 && ok; y--) {
        const float *src = rgba + (size_t)y * width * 4;
        for (int x = 0; x < width; x++) {
            row[(size_t)x * 3] = src[(size_t)x * 4];
            row[(size_t)x * 3 + 1] = src[(size_t)x * 4 + 1];
            row[(size_t)x * 3 + 2] = src[(size_t)x * 4 + 2];
        }
        ok = fwrite(row.data(), sizeof(float), row.size(), fp) == row.size();
    }
    if (fclose(fp) != 0) ok = false;
    return ok ? SRAY_OK : io_fail(SRAY_ERR_ARG, std::string("failed to write image: ") + path);
}

namespace {
void put_be32(std::vector<unsigned char> &v, uint32_t x)
{
    v.push_back((unsigned char)(x >> 24));
    v.push_back((unsigned char)(x >> 16));
    v.push_back((unsigned char)(x >> 8));
    v.push_back((unsigned char)x);
}

/* one PNG chunk: length, type, data, CRC-32 of type and data */
bool put_chunk(FILE *fp, const char type[4], const unsigned char *data, size_t n)
{
    std::vector<unsigned char> head;
    put_be32(head, (uint32_t)n);
    head.insert(head.end(), type, type + 4);
    uLong crc = crc32(0L, (const Bytef *)type, 4);
    if (n) crc = crc32(crc, data, (uInt)n);
    std::vector<unsigned char> tail;
    put_be32(tail, (uint32_t)crc);
    return fwrite(head.data(), 1, head.size(), fp) == head.size() && (n == 0 || fwrite(data, 1, n, fp) == n) &&
           fwrite(tail.data(), 1, tail.size(), fp) == tail.size();
}

/* a channel as a byte: value * 255 truncated toward zero, clamped to [0, 255] */
unsigned char to_byte(float f)
{
    const double v = (double)f * 255.0;
    if (!(v > 0.0)) return 0; /* negative, zero or NaN */
    if (v >= 255.0) return 255;
    return (unsigned char)(int)v;
}
} // namespace

/* The output image of the reference's main program (output.png or
 * frame%04d.png, src/sray.cc:150-155, saved by Image::save, src/img.cc:80-83,
 * through libimago, which converts the float image to 8-bit RGBA): a PNG of
 * colour type 6 (RGBA, 8 bits per channel), top row first, every scanline
 * with filter 0, one zlib stream in one IDAT chunk */
int sray_image_write_png(const char *path, const float *rgba, int width, int height)
{
    if (!path || !rgba || width <= 0 || height <= 0 || (uint64_t)width * (uint64_t)height > (1ull << 30))
        return io_fail(SRAY_ERR_ARG, "sray_image_write_png: bad arguments");
    const size_t stride = (size_t)width * 4 + 1;
    std::vector<unsigned char> raw, packed;
    try {
        raw.resize(stride * (size_t)height);
        packed.resize(compressBound((uLong)raw.size()));
    } catch (const std::bad_alloc &) {
        return io_fail(SRAY_ERR_NOMEM, "image too large for host memory");
    }
    for (int y = 0; y < height; y++) {
        unsigned char *row = &raw[(size_t)y * stride];
        row[0] = 0; /* filter type None */
        const float *src = rgba + (size_t)y * width * 4;
        for (size_t k = 0; k < (size_t)width * 4; k++) row[1 + k] = to_byte(src[k]);
    }
    uLongf plen = (uLongf)packed.size();
    if (compress2(packed.data(), &plen, raw.data(), (uLong)raw.size(), 6) != Z_OK)
        return io_fail(SRAY_ERR_ARG, "sray_image_write_png: compression failed");
    std::vector<unsigned char> ihdr;
    put_be32(ihdr, (uint32_t)width);
    put_be32(ihdr, (uint32_t)height);
    const unsigned char rest[5] = {8, 6, 0, 0, 0}; /* depth, RGBA, deflate, filter set 0, no interlace */
    ihdr.insert(ihdr.end(), rest, rest + 5);
    FILE *fp = fopen(path, "wb");
    if (!fp) return io_fail(SRAY_ERR_ARG, std::string("failed to open image file: ") + path + ": " + strerror(errno));
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    bool ok = fwrite(sig, 1, 8, fp) == 8 && put_chunk(fp, "IHDR", ihdr.data(), ihdr.size()) &&
              put_chunk(fp, "IDAT", packed.data(), plen) && put_chunk(fp, "IEND", nullptr, 0);
    if (fclose(fp) != 0) ok = false;
    return ok ? SRAY_OK : io_fail(SRAY_ERR_ARG, std::string("failed to write image: ") + path);
}

/* emit_status, src/render.cc:279-289: "%c %6d %6d %6d %6d\n" in a 32-byte
 * record, the rest zero */
int sray_status_record(char st, int a1, int a2, int a3, int a4, char *out)
{
    if (!out) return SRAY_ERR_ARG;
    char buf[64];
    int len = snprintf(buf, sizeof buf, "%c %6d %6d %6d %6d\n", st, a1, a2, a3, a4);
    if (len < 0 || len >= SRAY_STATUS_RECORD) return io_fail(SRAY_ERR_ARG, "status record longer than 31 bytes");
    memset(out, 0, SRAY_STATUS_RECORD);
    memcpy(out, buf, (size_t)len);
    return SRAY_OK;
}

} /* extern "C" */
