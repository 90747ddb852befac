/* PNG codec for the command-line pair evaluator (include/pdflow_png.h).  PNG structure per the
 * PNG specification (ISO/IEC 15948): signature, IHDR, PLTE, tRNS, IDAT (one zlib stream over all
 * rows, each row prefixed by its filter type 0-4), IEND; CRC-32 over type and data. */
#include "../../include/pdflow_png.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

static const uint8_t kSig[8] = {137, 80, 78, 71, 13, 10, 26, 10};

static uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

static void put32(uint8_t *p, uint32_t v)
{
    p[0] = (uint8_t)(v >> 24);
    p[1] = (uint8_t)(v >> 16);
    p[2] = (uint8_t)(v >> 8);
    p[3] = (uint8_t)v;
}

static int read_file(const char *path, uint8_t **buf, size_t *len)
{
    FILE *f = fopen(path, "rb");
    if (!f) return PDF_PNG_ERR_IO;
    if (fseek(f, 0, SEEK_END) != 0) { fclose(f); return PDF_PNG_ERR_IO; }
    long n = ftell(f);
    if (n < 0 || fseek(f, 0, SEEK_SET) != 0) { fclose(f); return PDF_PNG_ERR_IO; }
    *buf = (uint8_t *)malloc(n > 0 ? (size_t)n : 1);
    if (!*buf) { fclose(f); return PDF_PNG_ERR_NOMEM; }
    size_t got = fread(*buf, 1, (size_t)n, f);
    fclose(f);
    if (got != (size_t)n) { free(*buf); *buf = NULL; return PDF_PNG_ERR_IO; }
    *len = (size_t)n;
    return PDF_PNG_OK;
}

static int paeth(int a, int b, int c)
{
    const int p = a + b - c, pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

/* raw: h rows of (1 + stride) bytes; out: h * stride bytes */
static int unfilter(const uint8_t *raw, uint8_t *out, uint32_t h, size_t stride, unsigned bpp)
{
    for (uint32_t y = 0; y < h; y++) {
        const uint8_t ft = raw[y * (stride + 1)];
        const uint8_t *src = raw + y * (stride + 1) + 1;
        uint8_t *row = out + y * stride;
        const uint8_t *up = y ? out + (y - 1) * stride : NULL;
        for (size_t i = 0; i < stride; i++) {
            const int a = i >= bpp ? row[i - bpp] : 0, b = up ? up[i] : 0, c = (up && i >= bpp) ? up[i - bpp] : 0;
            int v;
            switch (ft) {
            case 0: v = src[i]; break;
            case 1: v = src[i] + a; break;
            case 2: v = src[i] + b; break;
            case 3: v = src[i] + ((a + b) >> 1); break;
            case 4: v = src[i] + paeth(a, b, c); break;
            default: return PDF_PNG_ERR_FORMAT;
            }
            row[i] = (uint8_t)v;
        }
    }
    return PDF_PNG_OK;
}

int pdf_png_read(const char *path, unsigned *width, unsigned *height, unsigned *channels,
                 unsigned *bit_depth, void **data)
{
    if (!path || !width || !height || !channels || !bit_depth || !data) return PDF_PNG_ERR_IO;
    *data = NULL;
    uint8_t *buf = NULL, *idat = NULL, *raw = NULL, *img = NULL;
    size_t len = 0, idat_len = 0;
    int rc = read_file(path, &buf, &len);
    if (rc) return rc;
    uint32_t w = 0, h = 0;
    unsigned depth = 0, ctype = 0, nch = 0, have_ihdr = 0, have_end = 0;
    uint8_t plte[768];
    unsigned nplte = 0, ntrns = 0;
    uint8_t trns[256];
    rc = PDF_PNG_ERR_FORMAT;
    if (len < 8 || memcmp(buf, kSig, 8) != 0) goto done;
    for (size_t p = 8; p + 12 <= len && !have_end;) {
        const uint32_t n = be32(buf + p);
        const uint8_t *type = buf + p + 4, *d = buf + p + 8;
        if (n > len - p - 12) goto done;
        const uint32_t crc = (uint32_t)crc32(crc32(0L, type, 4), d, n);
        if (crc != be32(d + n)) goto done;
        if (!memcmp(type, "IHDR", 4)) {
            if (n != 13 || have_ihdr) goto done;
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            if (w == 0 || h == 0 || d[10] != 0 || d[11] != 0) goto done;
            if (d[12] != 0) { rc = PDF_PNG_ERR_UNSUPPORTED; goto done; }
            switch (ctype) {
            case 0: nch = 1; break;
            case 2: nch = 3; break;
            case 3: nch = 1; break;
            case 4: nch = 2; break;
            case 6: nch = 4; break;
            default: goto done;
            }
            if (!(depth == 8 || (depth == 16 && ctype != 3) ||
                  ((depth == 1 || depth == 2 || depth == 4) && (ctype == 0 || ctype == 3)))) {
                rc = PDF_PNG_ERR_UNSUPPORTED;
                goto done;
            }
            have_ihdr = 1;
        } else if (!memcmp(type, "PLTE", 4)) {
            if (n % 3 || n > 768) goto done;
            memcpy(plte, d, n);
            nplte = n / 3;
        } else if (!memcmp(type, "tRNS", 4)) {
            if (ctype == 3) {
                if (n > 256) goto done;
                memcpy(trns, d, n);
                ntrns = n;
            }
        } else if (!memcmp(type, "IDAT", 4)) {
            uint8_t *q = (uint8_t *)realloc(idat, idat_len + n + 1);
            if (!q) { rc = PDF_PNG_ERR_NOMEM; goto done; }
            idat = q;
            memcpy(idat + idat_len, d, n);
            idat_len += n;
        } else if (!memcmp(type, "IEND", 4)) {
            have_end = 1;
        } else if (!(type[0] & 0x20)) {
            rc = PDF_PNG_ERR_UNSUPPORTED;     /* unknown critical chunk */
            goto done;
        }
        p += (size_t)n + 12;
    }
    if (!have_ihdr || !idat_len || (ctype == 3 && !nplte)) goto done;
    {
        const unsigned bpp = depth < 8 ? 1 : nch * depth / 8;       /* filter unit in bytes */
        const size_t stride = depth < 8 ? ((size_t)w * depth + 7) / 8 : (size_t)w * bpp;
        if (stride < w * (size_t)depth / 8 || (size_t)h > ((size_t)-1) / (stride + 1) || (size_t)w * h / w != h) {
            rc = PDF_PNG_ERR_NOMEM;
            goto done;
        }
        uLongf raw_len = (uLongf)((stride + 1) * h);
        raw = (uint8_t *)malloc(raw_len);
        img = (uint8_t *)malloc(stride * h);
        if (!raw || !img) { rc = PDF_PNG_ERR_NOMEM; goto done; }
        const uLongf want = raw_len;
        if (uncompress(raw, &raw_len, idat, (uLong)idat_len) != Z_OK || raw_len != want) goto done;
        if ((rc = unfilter(raw, img, h, stride, bpp)) != PDF_PNG_OK) goto done;
        if (depth < 8) {
            /* unpack 1/2/4-bit samples (most significant first) to one byte each; grey levels
             * are scaled to 0..255 */
            uint8_t *u = (uint8_t *)malloc((size_t)w * h);
            if (!u) { rc = PDF_PNG_ERR_NOMEM; goto done; }
            const unsigned per = 8 / depth, mask = (1u << depth) - 1;
            for (size_t y = 0; y < h; y++)
                for (size_t x = 0; x < w; x++) {
                    const unsigned byte = img[y * stride + x / per];
                    unsigned v = (byte >> (8 - depth * (1 + x % per))) & mask;
                    if (ctype == 0) v = v * 255u / mask;
                    u[y * w + x] = (uint8_t)v;
                }
            free(img);
            img = u;
            depth = 8;
        }
        if (ctype == 3) {
            const unsigned out_ch = ntrns ? 4 : 3;
            uint8_t *rgb = (uint8_t *)malloc((size_t)w * h * out_ch);
            if (!rgb) { rc = PDF_PNG_ERR_NOMEM; goto done; }
            for (size_t i = 0; i < (size_t)w * h; i++) {
                const unsigned k = img[i];
                if (k >= nplte) { free(rgb); rc = PDF_PNG_ERR_FORMAT; goto done; }
                memcpy(rgb + i * out_ch, plte + 3 * k, 3);
                if (ntrns) rgb[i * out_ch + 3] = k < ntrns ? trns[k] : 255;
            }
            free(img);
            img = rgb;
            nch = out_ch;
        } else if (depth == 16) {
            uint16_t *s = (uint16_t *)img;
            for (size_t i = 0; i < (size_t)w * h * nch; i++) {
                const uint8_t *b = img + 2 * i;
                s[i] = (uint16_t)(b[0] << 8 | b[1]);
            }
        }
    }
    *width = w;
    *height = h;
    *channels = nch;
    *bit_depth = depth;
    *data = img;
    img = NULL;
    rc = PDF_PNG_OK;
done:
    free(buf);
    free(idat);
    free(raw);
    free(img);
    return rc;
}

void pdf_png_free(void *data) { free(data); }

static int put_chunk(FILE *f, const char *type, const uint8_t *d, uint32_t n)
{
    uint8_t hdr[8], tail[4];
    put32(hdr, n);
    memcpy(hdr + 4, type, 4);
    static const uint8_t none = 0;     /* crc32 with a null buffer restarts the CRC */
    uint32_t crc = (uint32_t)crc32(crc32(0L, (const Bytef *)type, 4), n ? d : &none, n);
    put32(tail, crc);
    return fwrite(hdr, 1, 8, f) == 8 && (n == 0 || fwrite(d, 1, n, f) == n) && fwrite(tail, 1, 4, f) == 4;
}

int pdf_png_write(const char *path, unsigned width, unsigned height, unsigned channels,
                  unsigned bit_depth, const void *data)
{
    if (!path || !data || width == 0 || height == 0) return PDF_PNG_ERR_IO;
    if (!(channels == 1 || channels == 3) || !(bit_depth == 8 || bit_depth == 16)) return PDF_PNG_ERR_UNSUPPORTED;
    const size_t bpp = channels * bit_depth / 8, stride = (size_t)width * bpp;
    const size_t raw_len = (stride + 1) * height;
    uint8_t *raw = (uint8_t *)malloc(raw_len);
    uLongf zlen = compressBound((uLong)raw_len);
    uint8_t *z = (uint8_t *)malloc(zlen);
    if (!raw || !z) { free(raw); free(z); return PDF_PNG_ERR_NOMEM; }
    for (size_t y = 0; y < height; y++) {
        uint8_t *row = raw + y * (stride + 1);
        row[0] = 0;
        if (bit_depth == 8) {
            memcpy(row + 1, (const uint8_t *)data + y * stride, stride);
        } else {
            const uint16_t *s = (const uint16_t *)data + y * (size_t)width * channels;
            for (size_t i = 0; i < (size_t)width * channels; i++) {
                row[1 + 2 * i] = (uint8_t)(s[i] >> 8);
                row[2 + 2 * i] = (uint8_t)s[i];
            }
        }
    }
    int rc = compress2(z, &zlen, raw, (uLong)raw_len, 6) == Z_OK ? PDF_PNG_OK : PDF_PNG_ERR_NOMEM;
    free(raw);
    if (rc) { free(z); return rc; }
    FILE *f = fopen(path, "wb");
    if (!f) { free(z); return PDF_PNG_ERR_IO; }
    uint8_t ihdr[13];
    put32(ihdr, width);
    put32(ihdr + 4, height);
    ihdr[8] = (uint8_t)bit_depth;
    ihdr[9] = channels == 1 ? 0 : 2;
    ihdr[10] = ihdr[11] = ihdr[12] = 0;
    int ok = fwrite(kSig, 1, 8, f) == 8 && put_chunk(f, "IHDR", ihdr, 13) && put_chunk(f, "IDAT", z, (uint32_t)zlen) &&
             put_chunk(f, "IEND", NULL, 0);
    free(z);
    if (fclose(f) != 0) ok = 0;
    return ok ? PDF_PNG_OK : PDF_PNG_ERR_IO;
}

int pdf_png_size(const char *path, unsigned *width, unsigned *height)
{
    uint8_t hdr[24];
    FILE *f = fopen(path, "rb");
    if (!f) return PDF_PNG_ERR_IO;
    const size_t got = fread(hdr, 1, 24, f);
    fclose(f);
    if (got != 24 || memcmp(hdr, kSig, 8) != 0 || memcmp(hdr + 12, "IHDR", 4) != 0) return PDF_PNG_ERR_FORMAT;
    *width = be32(hdr + 16);
    *height = be32(hdr + 20);
    return PDF_PNG_OK;
}

/* 8-bit grey of pixel i: grey as is, 16-bit keeps the high byte, colour by BT.601 weights in
 * 14-bit fixed point with rounding (0.299 R + 0.587 G + 0.114 B) */
static uint8_t grey8(const void *data, unsigned ch, unsigned depth, size_t i)
{
    unsigned s[3];
    const unsigned n = ch >= 3 ? 3 : 1;
    for (unsigned k = 0; k < n; k++)
        s[k] = depth == 16 ? ((const uint16_t *)data)[i * ch + k] >> 8 : ((const uint8_t *)data)[i * ch + k];
    if (n == 1) return (uint8_t)s[0];
    return (uint8_t)((s[0] * 4899u + s[1] * 9617u + s[2] * 1868u + 8192u) >> 14);
}

int pdf_load_frame(const char *intensity_path, const char *depth_path, unsigned width,
                   unsigned height, float *I, float *Z)
{
    unsigned w, h, ch, depth;
    void *data;
    int rc;
    if (intensity_path && I) {
        if ((rc = pdf_png_read(intensity_path, &w, &h, &ch, &depth, &data))) return rc;
        if (w != width || h != height) { pdf_png_free(data); return PDF_PNG_ERR_FORMAT; }
        for (unsigned u = 0; u < width; u++)          /* scene_flow_impair.cpp:228-232 */
            for (unsigned v = 0; v < height; v++)
                I[v + u * height] = (float)grey8(data, ch, depth, (size_t)v * width + u);
        pdf_png_free(data);
    }
    if (depth_path && Z) {
        if ((rc = pdf_png_read(depth_path, &w, &h, &ch, &depth, &data))) return rc;
        if (w != width || h != height) { pdf_png_free(data); return PDF_PNG_ERR_FORMAT; }
        if (ch != 1) { pdf_png_free(data); return PDF_PNG_ERR_UNSUPPORTED; }
        for (unsigned v = 0; v < height; v++)          /* :242-245, convertTo(CV_32FC1, 1/5000) */
            for (unsigned u = 0; u < width; u++) {
                const size_t i = (size_t)v * width + u;
                const double s = depth == 16 ? ((const uint16_t *)data)[i] : ((const uint8_t *)data)[i];
                Z[v + u * height] = (float)(s * (1.0 / 5000.0));
            }
        pdf_png_free(data);
    }
    return PDF_PNG_OK;
}
