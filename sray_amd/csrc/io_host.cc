/*
 * io_host.cc — host-side input/output of the path: the photon-map text
 * dump and restore (PhotonMap::dump / restore, src/pmap.cc:180-251), the
 * framebuffer written as PFM or PNG (Image::save, src/img.cc:80-83), and the
 * 32-byte status records of the reference's backend protocol (emit_status,
 * src/render.cc:279-289), which the serving engine also writes.
 */
#include <cerrno>
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

extern "C" {

const char *sray_io_last_error(void) { return g_io_err.c_str(); }

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
