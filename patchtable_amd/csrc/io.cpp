// Image and NNF file formats of the PatchTable drop-in.
//
//   PFM  -- read_pfm_file / read_pfm_file3 / write_pfm_file / write_pfm_file3 (solver/pfm.cpp:15-116):
//           "PF" (3 channels) or "Pf" (1 channel), "w h", a scale whose sign gives the byte
//           order (negative: little endian), then float32 rows stored bottom row first.
//   PNG  -- load_color_image<float> (solver/array.h:1637-1684): 8-bit RGB, value * (1/255.0)
//           computed in double and stored as float; alpha dropped, grey expanded to RGB.
//           Decoded here with zlib (non-interlaced images, bit depths 1-16, all colour types).
//   save_color_image<double> (solver/array.h:1745-1827): an NNF is written as a 3-channel PFM
//           of its values converted to float.
#include "pt_internal.h"

#include <zlib.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

bool ends_with(const char *s, const char *suffix) {
    size_t n = strlen(s), m = strlen(suffix);
    return n >= m && strcmp(s + n - m, suffix) == 0;
}

struct File {
    FILE *f = nullptr;
    explicit File(const char *path, const char *mode) : f(fopen(path, mode)) {}
    ~File() {
        if (f) fclose(f);
    }
};

// ---------------------------------------------------------------------------------------------
// PFM
// ---------------------------------------------------------------------------------------------
int pfm_read(const char *path, int *h, int *w, int *c, std::vector<float> *data) {
    File F(path, "rb");
    if (!F.f) return pt_failf(PT_E_IO, "could not open %s", path);
    char tag[8] = {0};
    int ww = 0, hh = 0;
    double scale = 0;
    if (fscanf(F.f, "%7s", tag) != 1 || fscanf(F.f, "%d %d", &ww, &hh) != 2 || fscanf(F.f, "%lf", &scale) != 1)
        return pt_failf(PT_E_IO, "%s: bad PFM header", path);
    fgetc(F.f);   // the single whitespace byte that ends the header
    int channels;
    if (!strcmp(tag, "PF")) channels = 3;
    else if (!strcmp(tag, "Pf")) channels = 1;
    else return pt_failf(PT_E_IO, "%s: not a PFM file", path);
    if (ww <= 0 || hh <= 0 || (long)ww * hh > (1L << 31) / 4) return pt_failf(PT_E_IO, "%s: bad PFM size", path);
    *h = hh;
    *w = ww;
    *c = channels;
    if (!data) return PT_OK;
    const size_t n = (size_t)ww * hh * channels;
    std::vector<uint8_t> raw(n * 4);
    if (fread(raw.data(), 4, n, F.f) != n) return pt_failf(PT_E_IO, "%s: truncated PFM data", path);
    const bool little = scale < 0;
    data->resize(n);
    for (size_t i = 0; i < n; i++) {
        uint8_t *p = &raw[i * 4];
        uint32_t u = little ? (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24
                            : (uint32_t)p[3] | (uint32_t)p[2] << 8 | (uint32_t)p[1] << 16 | (uint32_t)p[0] << 24;
        memcpy(&(*data)[i], &u, 4);
    }
    return PT_OK;
}

// ---------------------------------------------------------------------------------------------
// PNG (8-bit RGB output)
// ---------------------------------------------------------------------------------------------
uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

int paeth(int a, int b, int c) {
    int p = a + b - c, pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

struct Png {
    int w = 0, h = 0, depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> palette;   // RGB triples
    std::vector<uint8_t> idat;
};

int png_parse(const char *path, Png &P, bool header_only) {
    File F(path, "rb");
    if (!F.f) return pt_failf(PT_E_IO, "could not open %s", path);
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    uint8_t hdr[8];
    if (fread(hdr, 1, 8, F.f) != 8 || memcmp(hdr, sig, 8)) return pt_failf(PT_E_IO, "%s: not a PNG file", path);
    for (;;) {
        uint8_t lt[8];
        if (fread(lt, 1, 8, F.f) != 8) return pt_failf(PT_E_IO, "%s: truncated PNG", path);
        const uint32_t len = be32(lt);
        const std::string type((const char *)lt + 4, 4);
        if (len > (1u << 30)) return pt_failf(PT_E_IO, "%s: bad PNG chunk", path);
        std::vector<uint8_t> d(len + 4);
        if (fread(d.data(), 1, len + 4, F.f) != len + 4) return pt_failf(PT_E_IO, "%s: truncated PNG chunk", path);
        const uint32_t crc = (uint32_t)crc32(crc32(0L, lt + 4, 4), d.data(), len);
        if (crc != be32(&d[len])) return pt_failf(PT_E_IO, "%s: PNG chunk %s fails its CRC", path, type.c_str());
        if (type == "IHDR") {
            if (len < 13) return pt_failf(PT_E_IO, "%s: bad IHDR", path);
            P.w = (int)be32(&d[0]);
            P.h = (int)be32(&d[4]);
            P.depth = d[8];
            P.ctype = d[9];
            P.interlace = d[12];
            if (header_only) return PT_OK;
        } else if (type == "PLTE") {
            P.palette.assign(d.begin(), d.begin() + len);
        } else if (type == "IDAT") {
            P.idat.insert(P.idat.end(), d.begin(), d.begin() + len);
        } else if (type == "IEND") {
            return PT_OK;
        }
    }
}

int png_decode(const char *path, int *h, int *w, std::vector<uint8_t> *rgb) {
    Png P;
    TRY_PT(png_parse(path, P, rgb == nullptr));
    if (P.w <= 0 || P.h <= 0 || (long)P.w * P.h > (1L << 30)) return pt_failf(PT_E_IO, "%s: bad PNG size", path);
    *h = P.h;
    *w = P.w;
    if (!rgb) return PT_OK;
    if (P.interlace) return pt_failf(PT_E_UNSUPPORTED, "%s: interlaced PNG is not supported", path);
    int spp;   // samples per pixel
    switch (P.ctype) {
    case 0: spp = 1; break;
    case 2: spp = 3; break;
    case 3: spp = 1; break;
    case 4: spp = 2; break;
    case 6: spp = 4; break;
    default: return pt_failf(PT_E_IO, "%s: bad PNG colour type %d", path, P.ctype);
    }
    const int bits = P.depth * spp;
    const size_t stride = ((size_t)P.w * bits + 7) / 8;
    const int bpp = (bits + 7) / 8;   // filter byte distance
    std::vector<uint8_t> raw((stride + 1) * P.h);
    z_stream zs;
    memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return pt_fail(PT_E_IO, "inflateInit failed");
    zs.next_in = P.idat.data();
    zs.avail_in = (uInt)P.idat.size();
    zs.next_out = raw.data();
    zs.avail_out = (uInt)raw.size();
    int zr = inflate(&zs, Z_FINISH);
    inflateEnd(&zs);
    if (zr != Z_STREAM_END || zs.avail_out != 0) return pt_failf(PT_E_IO, "%s: bad PNG image data", path);
    std::vector<uint8_t> cur(stride), prev(stride, 0);
    rgb->resize((size_t)P.w * P.h * 3);
    for (int y = 0; y < P.h; y++) {
        const uint8_t *in = &raw[(size_t)y * (stride + 1)];
        const int ft = in[0];
        for (size_t i = 0; i < stride; i++) {
            int a = i >= (size_t)bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= (size_t)bpp ? prev[i - bpp] : 0;
            int x = in[1 + i], v;
            switch (ft) {
            case 0: v = x; break;
            case 1: v = x + a; break;
            case 2: v = x + b; break;
            case 3: v = x + ((a + b) >> 1); break;
            case 4: v = x + paeth(a, b, c); break;
            default: return pt_failf(PT_E_IO, "%s: bad PNG filter %d", path, ft);
            }
            cur[i] = (uint8_t)v;
        }
        // samples -> 8-bit RGB (16-bit samples keep their high byte, low depths are scaled)
        auto sample = [&](int x, int s) -> int {
            const long bit = ((long)x * spp + s) * P.depth;
            if (P.depth == 8) return cur[bit >> 3];
            if (P.depth == 16) return cur[bit >> 3];
            const int shift = 8 - P.depth - (int)(bit & 7);
            return (cur[bit >> 3] >> shift) & ((1 << P.depth) - 1);
        };
        const int maxv = P.depth >= 8 ? 255 : (1 << P.depth) - 1;
        for (int x = 0; x < P.w; x++) {
            uint8_t *o = &(*rgb)[((size_t)y * P.w + x) * 3];
            if (P.ctype == 3) {
                const int idx = sample(x, 0);
                if ((size_t)idx * 3 + 2 >= P.palette.size()) return pt_failf(PT_E_IO, "%s: bad palette index", path);
                o[0] = P.palette[idx * 3];
                o[1] = P.palette[idx * 3 + 1];
                o[2] = P.palette[idx * 3 + 2];
            } else if (P.ctype == 0 || P.ctype == 4) {
                const int g = sample(x, 0) * 255 / maxv;
                o[0] = o[1] = o[2] = (uint8_t)g;
            } else {
                for (int ch = 0; ch < 3; ch++) o[ch] = (uint8_t)sample(x, ch);
            }
        }
        cur.swap(prev);
    }
    return PT_OK;
}

}  // namespace

extern "C" int pt_image_size(const char *path, int *h, int *w, int *c) {
    if (!path || !h || !w || !c) return pt_fail(PT_E_INVALID, "pt_image_size: null argument");
    if (ends_with(path, ".pfm")) return pfm_read(path, h, w, c, nullptr);
    *c = 3;
    return png_decode(path, h, w, nullptr);
}

extern "C" int pt_image_load(const char *path, float *out, int h, int w, int c) {
    if (!path || !out) return pt_fail(PT_E_INVALID, "pt_image_load: null argument");
    int hh = 0, ww = 0, cc = 0;
    if (ends_with(path, ".pfm")) {
        std::vector<float> d;
        TRY_PT(pfm_read(path, &hh, &ww, &cc, &d));
        if (hh != h || ww != w || cc != c) return pt_failf(PT_E_INVALID, "%s: expected %dx%dx%d", path, h, w, c);
        for (int y = 0; y < h; y++) {   // rows are stored bottom row first
            const float *src = &d[(size_t)(h - 1 - y) * w * c];
            memcpy(out + (size_t)y * w * c, src, (size_t)w * c * 4);
        }
        return PT_OK;
    }
    std::vector<uint8_t> rgb;
    TRY_PT(png_decode(path, &hh, &ww, &rgb));
    if (hh != h || ww != w || c != 3) return pt_failf(PT_E_INVALID, "%s: expected %dx%dx3", path, h, w);
    const double scale = 1.0 / 255.0;   // LOAD_SCALE_FACTOR(float), solver/array.h:1630
    for (size_t i = 0; i < rgb.size(); i++) out[i] = (float)(rgb[i] * scale);
    return PT_OK;
}

extern "C" int pt_pfm_save(const char *path, const double *data, int h, int w, int c) {
    if (!path || !data || h <= 0 || w <= 0 || (c != 1 && c != 3))
        return pt_fail(PT_E_INVALID, "pt_pfm_save: expected a 1 or 3 channel image");
    File F(path, "wb");
    if (!F.f) return pt_failf(PT_E_IO, "could not open %s for writing", path);
    fprintf(F.f, "%s\n%d %d\n%lf\n", c == 3 ? "PF" : "Pf", w, h, -1.0);
    std::vector<float> row((size_t)w * c);
    for (int y = h - 1; y >= 0; y--) {
        for (size_t i = 0; i < row.size(); i++) row[i] = (float)data[(size_t)y * w * c + i];
        if (fwrite(row.data(), 4, row.size(), F.f) != row.size()) return pt_failf(PT_E_IO, "%s: write failed", path);
    }
    return PT_OK;
}
