/* artful-hip — images and frame descriptions for the headless drivers (frames.h). */
#include "frames.h"

#include <stdlib.h>
#include <string.h>

static uint32_t crcTable[256];

static void crcInit(void) {
    for (uint32_t n = 0; n < 256; ++n) {
        uint32_t c = n;
        for (int k = 0; k < 8; ++k) c = c & 1 ? 0xEDB88320u ^ (c >> 1) : c >> 1;
        crcTable[n] = c;
    }
}

static uint32_t crcAdd(uint32_t crc, const unsigned char* p, size_t n) {
    for (size_t i = 0; i < n; ++i) crc = crcTable[(crc ^ p[i]) & 255] ^ (crc >> 8);
    return crc;
}

static void be32(unsigned char* p, uint32_t v) { p[0] = (unsigned char)(v >> 24), p[1] = (unsigned char)(v >> 16), p[2] = (unsigned char)(v >> 8), p[3] = (unsigned char)v; }

/* A PNG chunk: length, type, data, CRC of type and data. */
static bool chunk(FILE* out, const char* type, const unsigned char* data, uint32_t n) {
    unsigned char head[8];
    be32(head, n);
    memcpy(head + 4, type, 4);
    uint32_t crc = crcAdd(0xFFFFFFFFu, head + 4, 4);
    crc = crcAdd(crc, data, n) ^ 0xFFFFFFFFu;
    unsigned char tail[4];
    be32(tail, crc);
    return fwrite(head, 1, 8, out) == 8 && (n == 0 || fwrite(data, 1, n, out) == n) && fwrite(tail, 1, 4, out) == 4;
}

static bool savePng(FILE* out, const uint32_t* bgra, int width, int height) {
    crcInit();
    static const unsigned char signature[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    unsigned char header[13];
    be32(header, (uint32_t)width);
    be32(header + 4, (uint32_t)height);
    header[8] = 8, header[9] = 6, header[10] = 0, header[11] = 0, header[12] = 0; /* 8 bits, RGBA, deflate, no filter choice, no interlace */
    /* the image data: each row a filter byte (0) and RGBA, in zlib's stored blocks of at most 65535 bytes */
    const size_t rowBytes = (size_t)width * 4 + 1, raw = rowBytes * (size_t)height;
    const size_t blocks = raw / 65535 + 1;
    unsigned char* z = malloc(2 + raw + blocks * 5 + 4);
    unsigned char* row = malloc(rowBytes);
    size_t at = 0, left = raw, inBlock = 0;
    uint32_t a = 1, b = 0; /* Adler-32 of the raw bytes */
    z[at++] = 0x78, z[at++] = 0x01;
    for (int y = 0; y < height; ++y) {
        row[0] = 0;
        for (int x = 0; x < width; ++x) {
            const uint32_t p = bgra[(size_t)y * width + x];
            unsigned char* q = row + 1 + (size_t)x * 4;
            q[0] = (unsigned char)(p >> 16), q[1] = (unsigned char)(p >> 8), q[2] = (unsigned char)p, q[3] = (unsigned char)(p >> 24);
        }
        for (size_t i = 0; i < rowBytes; ++i) {
            if (inBlock == 0) { /* a stored block's header: last flag, then its length and that length's complement */
                const size_t n = left < 65535 ? left : 65535;
                z[at++] = left <= 65535 ? 1 : 0;
                z[at++] = (unsigned char)n, z[at++] = (unsigned char)(n >> 8);
                z[at++] = (unsigned char)~n, z[at++] = (unsigned char)(~n >> 8);
                inBlock = n;
            }
            z[at++] = row[i];
            a = (a + row[i]) % 65521u, b = (b + a) % 65521u;
            --inBlock, --left;
        }
    }
    be32(z + at, b << 16 | a);
    at += 4;
    const bool ok = fwrite(signature, 1, 8, out) == 8 && chunk(out, "IHDR", header, 13) && chunk(out, "IDAT", z, (uint32_t)at) && chunk(out, "IEND", NULL, 0);
    free(row);
    free(z);
    return ok;
}

static bool savePpm(FILE* out, const uint32_t* bgra, int width, int height) {
    fprintf(out, "P6\n%d %d\n255\n", width, height);
    unsigned char* row = malloc((size_t)width * 3);
    bool ok = true;
    for (int y = 0; y < height && ok; ++y) {
        for (int x = 0; x < width; ++x) {
            const uint32_t p = bgra[(size_t)y * width + x];
            row[x * 3] = (unsigned char)(p >> 16), row[x * 3 + 1] = (unsigned char)(p >> 8), row[x * 3 + 2] = (unsigned char)p;
        }
        ok = fwrite(row, 3, (size_t)width, out) == (size_t)width;
    }
    free(row);
    return ok;
}

bool driverSaveImage(const char* path, const uint32_t* bgra, int width, int height) {
    FILE* out = fopen(path, "wb");
    if (!out) return false;
    const size_t n = strlen(path);
    const bool ppm = n >= 4 && !strcmp(path + n - 4, ".ppm");
    const bool ok = ppm ? savePpm(out, bgra, width, height) : savePng(out, bgra, width, height);
    return fclose(out) == 0 && ok;
}

static void hex(FILE* out, const void* p, size_t n) {
    const unsigned char* b = p;
    for (size_t i = 0; i < n; ++i) fprintf(out, "%02x", b[i]);
}

void driverDescribe(FILE* out, int index, const ArtfBlackHoleDesc* desc, const ArtfCamera* camera) {
    if (!out) return;
    fprintf(out, "frame %d desc ", index);
    hex(out, desc, sizeof *desc);
    fputs(" camera ", out);
    hex(out, camera, sizeof *camera);
    fputc('\n', out);
    fflush(out);
}
