/*
 * gm_fits.cpp -- self-contained FITS codec for the gmilos command: primary
 * HDU float32 images (BITPIX -32, big-endian, 2880-byte records, 80-column
 * header cards), the formats the reference reads and writes with CFITSIO
 * (utilsFits.c:267-649, 2403-2520, 2777-2948).
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gmilos_io.h"

struct gmio_cube {
    int32_t naxis = 0;
    int64_t naxes[4] = {1, 1, 1, 1};
    int pos_lambda = -1, pos_row = -1, pos_col = -1, pos_stokes = -1;
    int32_t nrows = 0, ncols = 0, nlambda = 0;
    std::vector<std::string> cards;   /* input header cards (without END)    */
    std::vector<float> data;          /* [npix][4][nlambda]                   */
};

namespace {

constexpr int kBlock = 2880;

uint32_t bswap32(uint32_t x)
{
    return (x >> 24) | ((x >> 8) & 0xff00u) | ((x << 8) & 0xff0000u) | (x << 24);
}

std::string card_key(const std::string &c) { return c.substr(0, 8); }

std::string trim(const std::string &s)
{
    size_t a = s.find_first_not_of(" "), b = s.find_last_not_of(" ");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

/* Value of "KEY     = value / comment" (strings lose quotes and padding). */
bool card_value(const std::string &c, std::string &v)
{
    if (c.size() < 10 || c[8] != '=') return false;
    std::string rest = c.substr(10);
    rest = trim(rest);
    if (!rest.empty() && rest[0] == '\'') {
        size_t e = rest.find('\'', 1);
        v = trim(rest.substr(1, e == std::string::npos ? std::string::npos : e - 1));
        return true;
    }
    size_t slash = rest.find('/');
    v = trim(rest.substr(0, slash));
    return true;
}

std::string make_card(const char *key, const std::string &value, bool is_string = false)
{
    char buf[81];
    if (is_string) {
        std::string q = "'" + value;
        while (q.size() < 9) q += ' ';
        q += "'";
        std::snprintf(buf, sizeof buf, "%-8.8s= %-70s", key, q.c_str());
    } else {
        std::snprintf(buf, sizeof buf, "%-8.8s= %20s%50s", key, value.c_str(), "");
    }
    return std::string(buf, 80);
}

bool read_header(FILE *fp, std::vector<std::string> &cards)
{
    char block[kBlock];
    for (;;) {
        if (std::fread(block, 1, kBlock, fp) != (size_t)kBlock) return false;
        for (int i = 0; i < kBlock / 80; ++i) {
            std::string c(block + 80 * i, 80);
            if (card_key(c) == "END     ") return true;
            cards.push_back(c);
        }
    }
}

bool write_header(FILE *fp, const std::vector<std::string> &cards)
{
    std::string h;
    for (const auto &c : cards) h += c;
    h += std::string("END") + std::string(77, ' ');
    while (h.size() % kBlock) h += ' ';
    return std::fwrite(h.data(), 1, h.size(), fp) == h.size();
}

bool write_data(FILE *fp, const float *x, int64_t n)
{
    std::vector<uint32_t> buf(4096);
    int64_t done = 0;
    while (done < n) {
        const int64_t m = std::min<int64_t>(4096, n - done);
        for (int64_t i = 0; i < m; ++i) {
            uint32_t u;
            std::memcpy(&u, x + done + i, 4);
            buf[i] = bswap32(u);
        }
        if (std::fwrite(buf.data(), 4, m, fp) != (size_t)m) return false;
        done += m;
    }
    const int64_t bytes = n * 4;
    const int64_t pad = (kBlock - bytes % kBlock) % kBlock;
    std::vector<char> z(pad, 0);
    return pad == 0 || std::fwrite(z.data(), 1, pad, fp) == (size_t)pad;
}

struct ImageHeader {
    int bitpix = 0, naxis = 0;
    int64_t naxes[4] = {1, 1, 1, 1};
    std::string ctype[4];
    double bscale = 1.0, bzero = 0.0;
};

bool parse_header(const std::vector<std::string> &cards, ImageHeader &h)
{
    for (const auto &c : cards) {
        std::string k = trim(card_key(c)), v;
        if (!card_value(c, v)) continue;
        if (k == "BITPIX") h.bitpix = std::atoi(v.c_str());
        else if (k == "NAXIS") h.naxis = std::atoi(v.c_str());
        else if (k == "BSCALE") h.bscale = std::atof(v.c_str());
        else if (k == "BZERO") h.bzero = std::atof(v.c_str());
        else if (k.size() == 6 && k.compare(0, 5, "NAXIS") == 0 && k[5] >= '1' && k[5] <= '4')
            h.naxes[k[5] - '1'] = std::atoll(v.c_str());
        else if (k.size() == 6 && k.compare(0, 5, "CTYPE") == 0 && k[5] >= '1' && k[5] <= '4')
            h.ctype[k[5] - '1'] = v;
    }
    return h.naxis >= 1 && h.naxis <= 4;
}

bool read_floats(FILE *fp, const ImageHeader &h, int64_t n, std::vector<float> &out)
{
    out.resize(n);
    std::vector<uint32_t> buf(n);
    if (std::fread(buf.data(), 4, n, fp) != (size_t)n) return false;
    for (int64_t i = 0; i < n; ++i) {
        uint32_t u = bswap32(buf[i]);
        float f;
        std::memcpy(&f, &u, 4);
        out[i] = (h.bscale == 1.0 && h.bzero == 0.0) ? f : (float)(h.bzero + h.bscale * f);
    }
    return true;
}

/* Any FITS BITPIX (8, 16, 32, 64, -32, -64) converted to double with
 * BSCALE/BZERO, as CFITSIO's TDOUBLE reads do. */
bool read_doubles(FILE *fp, const ImageHeader &h, int64_t n, std::vector<double> &out)
{
    const int bytes = std::abs(h.bitpix) / 8;
    if (bytes != 1 && bytes != 2 && bytes != 4 && bytes != 8) return false;
    std::vector<unsigned char> raw((size_t)(n * bytes));
    if (std::fread(raw.data(), 1, raw.size(), fp) != raw.size()) return false;
    out.resize(n);
    for (int64_t i = 0; i < n; ++i) {
        const unsigned char *p = raw.data() + i * bytes;
        uint64_t u = 0;
        for (int b = 0; b < bytes; ++b) u = (u << 8) | p[b];
        double v;
        switch (h.bitpix) {
        case 8: v = (double)(uint8_t)u; break;
        case 16: v = (double)(int16_t)(uint16_t)u; break;
        case 32: v = (double)(int32_t)(uint32_t)u; break;
        case 64: v = (double)(int64_t)u; break;
        case -32: { uint32_t w = (uint32_t)u; float f; std::memcpy(&f, &w, 4); v = f; break; }
        case -64: std::memcpy(&v, &u, 8); break;
        default: return false;
        }
        out[i] = h.bzero + h.bscale * v;
    }
    return true;
}

} // namespace

/* Wavelength FITS file (readFitsLambdaToArray, utilsFits.c:1446-1529): a
 * 2-D image of naxes (nlambda, 2); the first row holds the line index (the
 * reference keeps the row's last value), the second the wavelengths in A.
 * The reference writes past its nlambda-sized array when the image has more
 * than two rows and reads into a null array for its 4-D form; both are
 * rejected here.  Returns nlambda, or a negative code. */
extern "C" int32_t gmio_fits_read_lambda(const char *path, int32_t *line_index, double *lambda,
                                         int32_t cap)
{
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return -1;
    std::vector<std::string> cards;
    ImageHeader h;
    std::vector<double> v;
    int32_t rc = -2;
    if (read_header(fp, cards) && parse_header(cards, h) && h.naxis == 2 && h.naxes[1] == 2 &&
        h.naxes[0] >= 1) {
        const int64_t n = h.naxes[0];
        if (n > cap) rc = -3;
        else if (read_doubles(fp, h, 2 * n, v)) {
            *line_index = (int32_t)v[n - 1];
            for (int64_t k = 0; k < n; ++k) lambda[k] = v[n + k];
            rc = (int32_t)n;
        }
    }
    std::fclose(fp);
    return rc;
}

extern "C" int32_t gmio_fits_write_image(const char *path, int32_t naxis, const int64_t *naxes,
                                         const char *const *ctypes, const float *data)
{
    if (naxis < 1 || naxis > 4) return -2;
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return -1;
    std::vector<std::string> cards;
    cards.push_back(make_card("SIMPLE", "T"));
    cards.push_back(make_card("BITPIX", "-32"));
    cards.push_back(make_card("NAXIS", std::to_string(naxis)));
    int64_t n = 1;
    for (int i = 0; i < naxis; ++i) {
        char k[9];
        std::snprintf(k, sizeof k, "NAXIS%d", i + 1);
        cards.push_back(make_card(k, std::to_string(naxes[i])));
        n *= naxes[i];
    }
    cards.push_back(make_card("EXTEND", "T"));
    if (ctypes)
        for (int i = 0; i < naxis; ++i)
            if (ctypes[i] && ctypes[i][0]) {
                char k[9];
                std::snprintf(k, sizeof k, "CTYPE%d", i + 1);
                cards.push_back(make_card(k, ctypes[i], true));
            }
    const bool ok = write_header(fp, cards) && write_data(fp, data, n);
    std::fclose(fp);
    return ok ? 0 : -1;
}

extern "C" int32_t gmio_fits_read_image(const char *path, int32_t *naxis, int64_t naxes[4],
                                        float *data, int64_t cap)
{
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return -1;
    std::vector<std::string> cards;
    ImageHeader h;
    if (!read_header(fp, cards) || !parse_header(cards, h) || h.bitpix != -32) {
        std::fclose(fp);
        return -2;
    }
    int64_t n = 1;
    for (int i = 0; i < h.naxis; ++i) n *= h.naxes[i];
    *naxis = h.naxis;
    for (int i = 0; i < 4; ++i) naxes[i] = h.naxes[i];
    if (data) {
        if (n > cap) { std::fclose(fp); return -3; }
        std::vector<float> v;
        if (!read_floats(fp, h, n, v)) { std::fclose(fp); return -2; }
        std::memcpy(data, v.data(), n * 4);
    }
    std::fclose(fp);
    return 0;
}

/* Axis roles: CTYPE WAVE-GRI / HPLN-TAN / HPLT-TAN / STOKES
 * (utilsFits.c:314-335); without CTYPEs, Stokes is the smallest axis and
 * wavelength the next smallest, the two remaining axes are row then column
 * (utilsFits.c:349-388). */
static bool axis_roles(const ImageHeader &h, gmio_cube &c)
{
    bool header = true;
    for (int i = 0; i < 4; ++i)
        if (h.ctype[i].empty()) header = false;
    if (header) {
        for (int i = 0; i < 4; ++i) {
            if (h.ctype[i] == "WAVE-GRI") c.pos_lambda = i;
            if (h.ctype[i] == "HPLN-TAN") c.pos_row = i;
            if (h.ctype[i] == "HPLT-TAN") c.pos_col = i;
            if (h.ctype[i] == "STOKES") c.pos_stokes = i;
        }
    } else {
        int64_t best = 10000;
        for (int i = 0; i < h.naxis; ++i)
            if (h.naxes[i] < best) { best = h.naxes[i]; c.pos_stokes = i; }
        best = 10000;
        for (int i = 0; i < h.naxis; ++i)
            if (i != c.pos_stokes && h.naxes[i] < best) { best = h.naxes[i]; c.pos_lambda = i; }
        int rest[2], k = 0;
        for (int i = 0; i < 4; ++i)
            if (i != c.pos_stokes && i != c.pos_lambda && k < 2) rest[k++] = i;
        c.pos_row = rest[0];
        c.pos_col = rest[1];
    }
    return c.pos_lambda >= 0 && c.pos_row >= 0 && c.pos_col >= 0 && c.pos_stokes >= 0;
}

extern "C" int32_t gmio_fits_read_cube(const char *path, int32_t nlambda_expected,
                                       gmio_cube **out)
{
    *out = nullptr;
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return -1;
    gmio_cube *c = new gmio_cube();
    ImageHeader h;
    if (!read_header(fp, c->cards) || !parse_header(c->cards, h) || h.bitpix != -32 ||
        h.naxis != 4 || !axis_roles(h, *c)) {
        std::fclose(fp);
        delete c;
        return -2;
    }
    c->naxis = h.naxis;
    for (int i = 0; i < 4; ++i) c->naxes[i] = h.naxes[i];
    c->nrows = (int32_t)h.naxes[c->pos_row];
    c->ncols = (int32_t)h.naxes[c->pos_col];
    c->nlambda = (int32_t)h.naxes[c->pos_lambda];
    if (h.naxes[c->pos_stokes] != 4 ||
        (nlambda_expected > 0 && c->nlambda != nlambda_expected)) {
        std::fclose(fp);
        delete c;
        return -3;
    }
    const int64_t n = h.naxes[0] * h.naxes[1] * h.naxes[2] * h.naxes[3];
    std::vector<float> raw;
    if (!read_floats(fp, h, n, raw)) {
        std::fclose(fp);
        delete c;
        return -2;
    }
    std::fclose(fp);
    c->data.assign(n, 0.f);
    const int nl = c->nlambda;
    int64_t idx[4];
    int64_t flat = 0;
    for (idx[3] = 0; idx[3] < h.naxes[3]; ++idx[3])
        for (idx[2] = 0; idx[2] < h.naxes[2]; ++idx[2])
            for (idx[1] = 0; idx[1] < h.naxes[1]; ++idx[1])
                for (idx[0] = 0; idx[0] < h.naxes[0]; ++idx[0], ++flat) {
                    const int64_t pix = idx[c->pos_col] * c->nrows + idx[c->pos_row];
                    c->data[(pix * 4 + idx[c->pos_stokes]) * nl + idx[c->pos_lambda]] = raw[flat];
                }
    *out = c;
    return 0;
}

extern "C" void gmio_cube_free(gmio_cube *c) { delete c; }

extern "C" int32_t gmio_cube_dims(const gmio_cube *c, int32_t *nrows, int32_t *ncols,
                                  int32_t *nlambda)
{
    if (!c) return -1;
    *nrows = c->nrows;
    *ncols = c->ncols;
    *nlambda = c->nlambda;
    return 0;
}

extern "C" const float *gmio_cube_data(const gmio_cube *c) { return c ? c->data.data() : nullptr; }

extern "C" int32_t gmio_fits_write_profiles(const char *path, const gmio_cube *like,
                                            const float *spectra)
{
    if (!like) return -2;
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return -1;
    const int64_t n = like->naxes[0] * like->naxes[1] * like->naxes[2] * like->naxes[3];
    std::vector<float> raw(n);
    const int nl = like->nlambda;
    int64_t idx[4];
    int64_t flat = 0;
    for (idx[3] = 0; idx[3] < like->naxes[3]; ++idx[3])
        for (idx[2] = 0; idx[2] < like->naxes[2]; ++idx[2])
            for (idx[1] = 0; idx[1] < like->naxes[1]; ++idx[1])
                for (idx[0] = 0; idx[0] < like->naxes[0]; ++idx[0], ++flat) {
                    const int64_t pix = idx[like->pos_col] * like->nrows + idx[like->pos_row];
                    raw[flat] = spectra[(pix * 4 + idx[like->pos_stokes]) * nl + idx[like->pos_lambda]];
                }
    std::vector<std::string> cards;
    for (const auto &cd : like->cards) {
        const std::string k = trim(card_key(cd));
        if (k == "BSCALE" || k == "BZERO") continue;   /* data written unscaled */
        cards.push_back(cd);
    }
    const bool ok = write_header(fp, cards) && write_data(fp, raw.data(), n);
    std::fclose(fp);
    return ok ? 0 : -1;
}

extern "C" int32_t gmio_fits_write_models(const char *path, int32_t nrows, int32_t ncols,
                                          const double *models, const float *chisqr,
                                          const int32_t *iters, int32_t with_chisqr)
{
    const int planes = 12 + (with_chisqr ? 1 : 0);
    const int64_t npix = (int64_t)nrows * ncols;
    std::vector<float> v((size_t)(npix * planes));
    for (int f = 0; f < planes; ++f)
        for (int64_t p = 0; p < npix; ++p) {
            float x;
            if (f < 11) x = (float)models[p * 11 + f];
            else if (f == 11) x = (float)iters[p];
            else x = chisqr[p];
            v[f * npix + p] = x;
        }
    const int64_t naxes[3] = {nrows, ncols, planes};
    std::remove(path);
    return gmio_fits_write_image(path, 3, naxes, nullptr, v.data());
}
