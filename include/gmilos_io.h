/*
 * gmilos_io.h -- file surface of the gmilos command (host only, no device).
 *
 * The reference reads a positional control file (.mtrol), LINES, a
 * wavelength grid (.grid), models (.mod), PSF tables (.psf), single-pixel
 * profiles (.per) and FITS cubes through CFITSIO, and writes models and
 * synthetic profiles (readConfig.c, utilsFits.c, gmilos.cu).  These entry
 * points read and write the same formats; the FITS codec is self-contained
 * (no CFITSIO) and handles the float images the reference accepts.
 */
#ifndef GMILOS_IO_H
#define GMILOS_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GMIO_PATH 4096

/* Control file contents (ConfigControl, defines.h:215-269, the fields the
 * positional parser of readConfig.c:1421-2094 fills). */
typedef struct {
    int32_t number_of_cycles;
    char    observed_profiles[GMIO_PATH];
    char    stray_light_file[GMIO_PATH];
    char    psf_file[GMIO_PATH];
    char    wavelength_file[GMIO_PATH];
    char    atomic_parameters_file[GMIO_PATH];
    char    abundances_file[GMIO_PATH];
    char    initial_guess_model[GMIO_PATH];
    char    initial_guess_model_2[GMIO_PATH];
    double  weight[4];
    int32_t automatic_select_of_nodes;
    int32_t fix[11];                    /* eta0,B,vlos,dopp,aa,gm,az,S0,S1,mac,alfa */
    int32_t fix2[11];
    int32_t invert_filling_factor;
    double  mu;
    int32_t estimated_sn_for_i;
    int32_t continuum_contrast;
    double  initial_diagonal_element;
    int32_t use_fft;
    int32_t logclambda;
    int32_t num_streams;
    int32_t device_id;
    /* derived */
    double  noise;                      /* (1/SN)^2, default 1e-6            */
    double  toplim;
    int32_t convolve_with_psf;
    double  fwhm;                       /* > 0 when the PSF field is a number */
    int32_t use_malla_grid;             /* wavelength file is a .grid         */
    int32_t save_chisqr;
    int32_t save_synthesis_adjusted;
} gmio_control;

/* Returns 0 on success, -1 if the file cannot be read, -2 on a field that
 * cannot be parsed (the reference returns 0 from readTrolFile). */
int32_t gmio_read_mtrol(const char *path, gmio_control *c);
void    gmio_control_defaults(gmio_control *c);
/* .mod: "name: value" lines matched by substring (readConfig.c:495-550). */
int32_t gmio_read_mod(const char *path, double model[11]);
int32_t gmio_write_mod(const char *path, const double model[11], int32_t iters,
                       double chisqr);
/* .grid (readConfig.c:626-667): returns the line index, fills first/step/
 * last offsets in mA. */
int32_t gmio_read_grid(const char *path, double *first, double *step, double *last);
/* .psf table: two columns (offset mA, value).  Returns the number of rows
 * read (<= cap), -1 on error. */
int32_t gmio_read_psf(const char *path, double *dl, double *val, int32_t cap);
int32_t gmio_count_lines(const char *path);
/* .per: 6 columns (index, offset, I, Q, U, V); spectro [4][nlambda]. */
int32_t gmio_read_per(const char *path, int32_t nlambda, float *spectro);
int32_t gmio_write_per(const char *path, int32_t line_index, int32_t nlambda,
                       const double *offsets_mA, const float *spectro);
/* Stray-light .per (readPerStrayLightFile, utilsFits.c:1532-1575): as
 * gmio_read_per, and the offsets column must equal trunc(offsets_mA[k]);
 * -3 when it does not. */
int32_t gmio_read_stray_per(const char *path, int32_t nlambda, const double *offsets_mA,
                            float *spectro);

/* ---- FITS ---------------------------------------------------------------- */
typedef struct gmio_cube gmio_cube;
/* Reads a float32 4-D Stokes cube (utilsFits.c:267-649): axes from CTYPE
 * (WAVE-GRI, HPLN-TAN, HPLT-TAN, STOKES) or, without them, Stokes = the
 * smallest axis and wavelength = the next smallest.  Pixels are ordered
 * col * nrows + row with spectra [4][nlambda] each. */
int32_t gmio_fits_read_cube(const char *path, int32_t nlambda_expected, gmio_cube **cube);
void    gmio_cube_free(gmio_cube *cube);
int32_t gmio_cube_dims(const gmio_cube *cube, int32_t *nrows, int32_t *ncols,
                       int32_t *nlambda);
const float *gmio_cube_data(const gmio_cube *cube);   /* [npix][4][nlambda] */
/* Writes spectra [npix][4][nlambda] with the input cube's header and axis
 * order (writeFitsImageProfiles, utilsFits.c:2777-2948). */
int32_t gmio_fits_write_profiles(const char *path, const gmio_cube *like,
                                 const float *spectra);
/* Model image (nrows, ncols, 12 or 13 planes: 11 model fields, iterations,
 * chi^2) as writeFitsImageModels (utilsFits.c:2403-2520). */
int32_t gmio_fits_write_models(const char *path, int32_t nrows, int32_t ncols,
                               const double *models, const float *chisqr,
                               const int32_t *iters, int32_t with_chisqr);
/* Generic float32 image writer/reader used by the tests and tools. */
int32_t gmio_fits_write_image(const char *path, int32_t naxis, const int64_t *naxes,
                              const char *const *ctypes, const float *data);
int32_t gmio_fits_read_image(const char *path, int32_t *naxis, int64_t naxes[4],
                             float *data, int64_t cap);
/* Wavelength FITS file (readFitsLambdaToArray, utilsFits.c:1446-1529):
 * naxes (nlambda, 2), row 1 the line index, row 2 the wavelengths in A; any
 * BITPIX.  Returns nlambda (<= cap) or a negative code. */
int32_t gmio_fits_read_lambda(const char *path, int32_t *line_index, double *lambda,
                              int32_t cap);

/* 1 when the file can be opened for reading (access(F_OK) of the reference) */
int32_t gmio_file_readable(const char *path);
/* file_ext / get_basefilename of readConfig.c:2312-2349 */
const char *gmio_file_ext(const char *path);
void gmio_base_name(const char *path, char *out, int32_t cap);

#ifdef __cplusplus
}
#endif
#endif
