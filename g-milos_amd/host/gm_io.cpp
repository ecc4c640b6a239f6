/*
 * gm_io.cpp -- text formats of the gmilos command: .mtrol, .mod, .grid, .psf,
 * .per (see gmilos_io.h).  Field layouts follow the reference's readers so
 * its example files are read unchanged.
 */
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unistd.h>

#include "gmilos_io.h"

extern "C" void gmio_control_defaults(gmio_control *c)
{
    std::memset(c, 0, sizeof *c);
    for (int i = 0; i < 4; ++i) c->weight[i] = 1.0;        /* readConfig.c:707-710 */
    c->noise = 1e-3 * 1e-3;                                /* NOISE_SIGMA^2        */
    c->initial_diagonal_element = 0.1;                     /* ILAMBDA              */
    c->toplim = 1e-12;                                     /* TOPLIM               */
    c->mu = 1.0;                                           /* AH                   */
    c->save_chisqr = 1;
    c->save_synthesis_adjusted = 1;
    c->num_streams = 2;
    c->fwhm = -1;
}

extern "C" const char *gmio_file_ext(const char *s)
{
    const char *ext = std::strrchr(s, '.');
    if (!ext) return s + std::strlen(s);
    for (const char *p = ext + 1; *p; ++p)
        if (!std::isalnum((unsigned char)*p)) return s + std::strlen(s);
    return ext;
}

extern "C" void gmio_base_name(const char *path, char *out, int32_t cap)
{
    /* the path up to its last '.', as get_basefilename (readConfig.c:2331-2349) */
    const char *dot = std::strrchr(path, '.');
    size_t n = dot ? (size_t)(dot - path) : std::strlen(path);
    if (n >= (size_t)cap) n = cap - 1;
    std::memcpy(out, path, n);
    out[n] = 0;
}

namespace {

/* One "name : value ! comment" line: everything after the first ':' up to
 * '!' is the value field (readConfig.c uses "%99[^:]:%s%99[^!]!").       */
bool value_field(const char *line, std::string &val)
{
    const char *colon = std::strchr(line, ':');
    if (!colon) return false;
    const char *p = colon + 1;
    while (*p == ' ' || *p == '\t') ++p;
    const char *e = p;
    while (*e && *e != '!' && *e != '\n' && *e != '\r' && *e != ' ' && *e != '\t') ++e;
    val.assign(p, e - p);
    return true;
}

bool int_field(const std::string &v, int32_t *out)
{
    if (v.empty()) return false;                 /* blank keeps the default */
    char *end = nullptr;
    long x = std::strtol(v.c_str(), &end, 0);
    if (end == v.c_str()) return false;
    *out = (int32_t)x;
    return true;
}

bool real_field(const std::string &v, double *out)
{
    if (v.empty()) return false;
    char *end = nullptr;
    double x = std::strtod(v.c_str(), &end);
    if (end == v.c_str()) return false;
    *out = (double)(float)x;                     /* the reference reads REAL */
    return true;
}

void str_field(const std::string &v, char *out)
{
    std::snprintf(out, GMIO_PATH, "%s", v.c_str());
}

} // namespace

/* Positional control file, readConfig.c:1421-2094 (44 lines in order). */
extern "C" int32_t gmio_read_mtrol(const char *path, gmio_control *c)
{
    gmio_control_defaults(c);
    FILE *fp = std::fopen(path, "r");
    if (!fp) return -1;
    char line[4096];
    std::string v;
    int idx = 0;
    int32_t *fixmap[10] = {&c->fix[7], &c->fix[8], &c->fix[0], &c->fix[1], &c->fix[2],
                           &c->fix[5], &c->fix[6], &c->fix[3], &c->fix[4], &c->fix[9]};
    int32_t *fix2map[10] = {&c->fix2[7], &c->fix2[8], &c->fix2[0], &c->fix2[1], &c->fix2[2],
                            &c->fix2[5], &c->fix2[6], &c->fix2[3], &c->fix2[4], &c->fix2[9]};
    double last_real = 1.0;
    int32_t rc = 0;
    while (idx < 44 && std::fgets(line, sizeof line, fp)) {
        if (!value_field(line, v)) { rc = -2; break; }
        const int i = idx++;
        double r;
        switch (i) {
        case 0: if (!int_field(v, &c->number_of_cycles)) rc = -2; break;
        case 1: str_field(v, c->observed_profiles); break;
        case 2: str_field(v, c->stray_light_file); break;
        case 3: {
            str_field(v, c->psf_file);
            if (v.empty() || access(c->psf_file, F_OK) != 0) {
                const double fwhm = std::atof(v.c_str());
                if (fwhm != 0) { c->fwhm = fwhm; c->convolve_with_psf = 1; }
                else { c->convolve_with_psf = 0; c->fwhm = -1; }
            } else {
                if (std::strcmp(gmio_file_ext(c->psf_file), ".psf") != 0) rc = -2;
                c->convolve_with_psf = 1;
            }
            break;
        }
        case 4: {
            str_field(v, c->wavelength_file);
            const char *e = gmio_file_ext(c->wavelength_file);
            c->use_malla_grid = std::strcmp(e, ".grid") == 0;
            break;
        }
        case 5: str_field(v, c->atomic_parameters_file); break;
        case 6: str_field(v, c->abundances_file); break;
        case 7: str_field(v, c->initial_guess_model); break;
        case 8: str_field(v, c->initial_guess_model_2); break;
        case 9: case 10: case 11: case 12:
            /* a blank weight keeps the value parsed last (the reference
             * reuses its scratch variable) */
            if (real_field(v, &r)) last_real = r;
            c->weight[i - 9] = last_real;
            break;
        case 13: int_field(v, &c->automatic_select_of_nodes); break;
        case 14: case 15: case 16: case 17: case 18: case 19: case 20: case 21: case 22:
        case 23:
            int_field(v, fixmap[i - 14]);
            break;
        case 24: case 25: case 26: case 27: case 28: case 29: case 30: case 31: case 32:
        case 33:
            int_field(v, fix2map[i - 24]);
            break;
        case 34: int_field(v, &c->invert_filling_factor); break;
        case 35: int_field(v, &c->fix[10]); c->fix2[10] = c->fix[10]; break;
        case 36: real_field(v, &c->mu); break;
        case 37:
            if (int_field(v, &c->estimated_sn_for_i) && c->estimated_sn_for_i != 0) {
                const double n = 1.0 / c->estimated_sn_for_i;
                c->noise = (double)(float)(n * n);
            }
            break;
        case 38: int_field(v, &c->continuum_contrast); break;
        case 39:
            if (real_field(v, &r)) last_real = r;
            c->initial_diagonal_element = last_real;
            break;
        case 40: int_field(v, &c->use_fft); break;
        case 41: int_field(v, &c->logclambda); break;
        case 42: int_field(v, &c->num_streams); break;
        case 43: int_field(v, &c->device_id); break;
        }
        if (rc) break;
    }
    std::fclose(fp);
    if (!rc && idx < 44) rc = -2;
    /* "!" as the first character of a value means empty (readConfig.c:1460) */
    char *strs[] = {c->observed_profiles, c->stray_light_file, c->psf_file,
                    c->wavelength_file, c->atomic_parameters_file, c->abundances_file,
                    c->initial_guess_model, c->initial_guess_model_2};
    for (char *s : strs)
        if (s[0] == '!') s[0] = 0;
    return rc;
}

/* Substring keys as readConfig.c:514-546. */
extern "C" int32_t gmio_read_mod(const char *path, double m[11])
{
    FILE *fp = std::fopen(path, "r");
    if (!fp) return -1;
    static const char *keys[11] = {"eta_0", "magnetic field", "LOS velocity", "Doppler width",
                                   "damping", "gamma", "phi", "S_0", "S_1", "v_mac",
                                   "filling factor"};
    char line[1024];
    while (std::fgets(line, sizeof line, fp)) {
        const char *colon = std::strchr(line, ':');
        if (!colon) continue;
        std::string name(line, colon - line);
        char *end = nullptr;
        const double val = std::strtod(colon + 1, &end);
        if (end == colon + 1) continue;
        for (int f = 0; f < 11; ++f)
            if (name.find(keys[f]) != std::string::npos) m[f] = val;
    }
    std::fclose(fp);
    return 0;
}

/* Output model file as gmilos.cu:844-860. */
extern "C" int32_t gmio_write_mod(const char *path, const double m[11], int32_t iters,
                                  double chisqr)
{
    FILE *fp = std::fopen(path, "w");
    if (!fp) return -1;
    std::fprintf(fp, "eta_0               :%lf\n", m[0]);
    std::fprintf(fp, "magnetic field [G]  :%lf\n", m[1]);
    std::fprintf(fp, "LOS velocity[km/s]  :%lf\n", m[2]);
    std::fprintf(fp, "Doppler width [A]   :%lf\n", m[3]);
    std::fprintf(fp, "damping             :%lf\n", m[4]);
    std::fprintf(fp, "gamma [deg]         :%lf\n", m[5]);
    std::fprintf(fp, "phi  [deg]          :%lf\n", m[6]);
    std::fprintf(fp, "S_0                 :%lf\n", m[7]);
    std::fprintf(fp, "S_1                 :%lf\n", m[8]);
    std::fprintf(fp, "v_mac [km/s]        :%lf\n", m[9]);
    std::fprintf(fp, "filling factor      :%lf\n", m[10]);
    std::fprintf(fp, "# Iterations        :%d\n", iters);
    std::fprintf(fp, "chisqr              :%le\n", chisqr);
    std::fprintf(fp, "\n\n");
    std::fclose(fp);
    return 0;
}

/* Grid: the data line follows the first line starting with "------"
 * ("%i%99[^:]:%lf,%lf,%lf", readConfig.c:643-655). */
extern "C" int32_t gmio_read_grid(const char *path, double *first, double *step, double *last)
{
    FILE *fp = std::fopen(path, "r");
    if (!fp) return 0;
    char line[1024];
    bool found = false;
    int32_t index = 0;
    while (std::fgets(line, sizeof line, fp)) {
        if (found) {
            char name[100];
            if (std::sscanf(line, "%i%99[^:]:%lf,%lf,%lf", &index, name, first, step, last) == 5) {
                std::fclose(fp);
                return index;
            }
            index = 0;
            break;
        }
        if (std::strncmp(line, "------", 6) == 0) found = true;
    }
    std::fclose(fp);
    return 0;
}

extern "C" int32_t gmio_count_lines(const char *path)
{
    FILE *fp = std::fopen(path, "r");
    if (!fp) return -1;
    int32_t n = 0;
    int ch;
    while ((ch = std::fgetc(fp)) != EOF)
        if (ch == '\n') ++n;
    std::fclose(fp);
    return n;
}

/* "%le %le" per line (readConfig.c:690-697). */
extern "C" int32_t gmio_read_psf(const char *path, double *dl, double *val, int32_t cap)
{
    FILE *fp = std::fopen(path, "r");
    if (!fp) return -1;
    char line[256];
    int32_t n = 0;
    while (n < cap && std::fgets(line, sizeof line, fp)) {
        double a, b;
        if (std::sscanf(line, "%le %le", &a, &b) == 2) {
            dl[n] = a;
            val[n] = b;
            ++n;
        }
    }
    std::fclose(fp);
    return n;
}

/* .per: "%e %e %e %e %e %e", columns 3-6 = I, Q, U, V (gmilos.cu:669-676). */
extern "C" int32_t gmio_read_per(const char *path, int32_t nlambda, float *sp)
{
    FILE *fp = std::fopen(path, "r");
    if (!fp) return -1;
    char line[512];
    int32_t k = 0;
    while (k < nlambda && std::fgets(line, sizeof line, fp)) {
        float a[6];
        if (std::sscanf(line, "%e %e %e %e %e %e", &a[0], &a[1], &a[2], &a[3], &a[4], &a[5]) != 6)
            continue;
        for (int s = 0; s < 4; ++s) sp[s * nlambda + k] = a[2 + s];
        ++k;
    }
    std::fclose(fp);
    return k == nlambda ? 0 : -2;
}

extern "C" int32_t gmio_file_readable(const char *path)
{
    if (!path || !path[0]) return 0;
    FILE *fp = std::fopen(path, "r");
    if (!fp) return 0;
    std::fclose(fp);
    return 1;
}

/* Stray-light profile in .per form (readPerStrayLightFile,
 * utilsFits.c:1532-1575): the second column must equal the truncated grid
 * offset of each row.  Returns 0, -1 (no file), -2 (short file) or -3
 * (offsets differ, where the reference exits).                              */
extern "C" int32_t gmio_read_stray_per(const char *path, int32_t nlambda, const double *off,
                                       float *sp)
{
    FILE *fp = std::fopen(path, "r");
    if (!fp) return -1;
    char line[512];
    int32_t k = 0, rc = 0;
    while (k < nlambda && std::fgets(line, sizeof line, fp)) {
        float a[6];
        if (std::sscanf(line, "%e %e %e %e %e %e", &a[0], &a[1], &a[2], &a[3], &a[4], &a[5]) != 6)
            continue;
        if (a[1] != (float)std::trunc(off[k])) rc = -3;
        for (int s = 0; s < 4; ++s) sp[s * nlambda + k] = a[2 + s];
        ++k;
    }
    std::fclose(fp);
    if (rc) return rc;
    return k == nlambda ? 0 : -2;
}

/* "%d\t%f\t%e\t%e\t%e\t%e\n" per wavelength (gmilos.cu:589, 885). */
extern "C" int32_t gmio_write_per(const char *path, int32_t line_index, int32_t nlambda,
                                  const double *off, const float *sp)
{
    FILE *fp = std::fopen(path, "w");
    if (!fp) return -1;
    for (int k = 0; k < nlambda; ++k)
        std::fprintf(fp, "%d\t%f\t%e\t%e\t%e\t%e\n", line_index, off[k], sp[k], sp[nlambda + k],
                     sp[2 * nlambda + k], sp[3 * nlambda + k]);
    std::fclose(fp);
    return 0;
}
