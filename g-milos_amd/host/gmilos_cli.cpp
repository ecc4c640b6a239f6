/*
 * gmilos_cli.cpp -- the `gmilos <file>.mtrol` command over the C ABI.
 *
 * Same file surface and modes as the reference's main (gmilos.cu:86-1341):
 *   NumberOfCycles <  0  classical estimates  -> <base>_model_ce.mod / .fits
 *   NumberOfCycles == 0  synthesis of the initial model -> <base>.per
 *   NumberOfCycles >  0  inversion of a .per pixel -> <base>_model.mod and
 *                        <base>_stokes.per, or of a FITS cube ->
 *                        <base>_mod.fits and <base>_stokes.fits
 * Options: --mixed (FP64 solve and chi^2 test), --ldl (LDL^T solve),
 * --cordic (the reference's 18-step CORDIC step solve), --jacobi (every step
 * by the Jacobi SVD), --exact-rf (exact magneto-optical response functions).
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gmilos.h"
#include "gmilos_io.h"

namespace {

int fail(const char *msg, const char *arg = "")
{
    std::fprintf(stderr, "gmilos: %s%s\n", msg, arg);
    return EXIT_FAILURE;
}

struct Dev {
    void *p = nullptr;
    explicit Dev(size_t bytes) { if (hipMalloc(&p, bytes ? bytes : 4) != hipSuccess) p = nullptr; }
    ~Dev() { if (p) (void)hipFree(p); }
};

std::string base_of(const gmio_control &c)
{
    char buf[GMIO_PATH];
    gmio_base_name(c.observed_profiles[0] ? c.observed_profiles : c.initial_guess_model, buf,
                   sizeof buf);
    return buf;
}

} // namespace

int main(int argc, char **argv)
{
    if (argc < 2) return fail("usage: gmilos <file.mtrol> [--mixed] [--ldl] [--cordic] [--jacobi] [--exact-rf]");
    int precision = GM_PREC_FP32, solver = GM_SOLVE_SVD, exact_rf = 0;
    for (int i = 2; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--mixed")) precision = GM_PREC_MIXED;
        else if (!std::strcmp(argv[i], "--ldl")) solver = GM_SOLVE_LDL;
        else if (!std::strcmp(argv[i], "--cordic")) solver = GM_SOLVE_CORDIC;   /* the reference's solve */
        else if (!std::strcmp(argv[i], "--jacobi")) solver = GM_SOLVE_JACOBI;   /* every step by the SVD */
        else if (!std::strcmp(argv[i], "--exact-rf")) exact_rf = 1;
        else return fail("unknown option ", argv[i]);
    }
    gmio_control cfg;
    if (gmio_read_mtrol(argv[1], &cfg) != 0) return fail("cannot parse control file ", argv[1]);
    if (hipSetDevice(cfg.device_id) != hipSuccess) return fail("cannot select the device");

    double init[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1.0};
    if (cfg.initial_guess_model[0] && gmio_read_mod(cfg.initial_guess_model, init) != 0)
        return fail("cannot read initial model ", cfg.initial_guess_model);
    /* gmilos.cu:163-166 */
    if (init[10] < 1 && !gmio_file_readable(cfg.stray_light_file))
        return fail("filling factor in the initial model is less than 1 and the stray light "
                    "file cannot be accessed: ", cfg.stray_light_file);

    /* wavelengths: .grid offsets (gmilos.cu:171-196) or a FITS file of
     * absolute wavelengths (gmilos.cu:197-209) */
    int index = 0, n = 0;
    double step = 0, lambda0, qn[6];
    std::vector<double> off, lam;
    if (cfg.use_malla_grid) {
        double first, last;
        index = gmio_read_grid(cfg.wavelength_file, &first, &step, &last);
        if (!index) return fail("cannot read grid ", cfg.wavelength_file);
        n = (int)((last - first) / step) + 1;
        if (n < 2 || n > GM_MAX_LAMBDA) return fail("grid size out of range");
        off.resize(n);
        lam.resize(n);
        off[0] = first;
        for (int k = 1; k < n; ++k) off[k] = off[k - 1] + step;
        if (gm_parse_lines_file(cfg.atomic_parameters_file, index, &lambda0, qn) != GM_OK)
            return fail("quantum line not found in ", cfg.atomic_parameters_file);
        lam[0] = lambda0 + first / 1000.0;
        for (int k = 1; k < n; ++k) lam[k] = lam[k - 1] + step / 1000.0;
    } else if (!std::strcmp(gmio_file_ext(cfg.wavelength_file), ".fits")) {
        lam.resize(GM_MAX_LAMBDA);
        n = gmio_fits_read_lambda(cfg.wavelength_file, &index, lam.data(), GM_MAX_LAMBDA);
        if (n < 2) return fail("the wavelength file has not been read properly: ", cfg.wavelength_file);
        lam.resize(n);
        if (gm_parse_lines_file(cfg.atomic_parameters_file, index, &lambda0, qn) != GM_OK)
            return fail("quantum line not found in ", cfg.atomic_parameters_file);
        /* the reference leaves the mA offsets unset on this path; a PSF
         * table is resampled here at the offsets from the line centre */
        off.resize(n);
        for (int k = 0; k < n; ++k) off[k] = (lam[k] - lambda0) * 1000.0;
        step = off[1] - off[0];
    } else {
        return fail("the wavelength file must be .grid or .fits: ", cfg.wavelength_file);
    }

    gm_setup s;
    gm_setup_defaults(&s);
    s.nlambda = n;
    for (int k = 0; k < n; ++k) s.lambda[k] = lam[k];
    s.nlines = 1;
    s.lambda0[0] = lambda0;
    if (gm_zeeman_from_quantum(qn[0], qn[1], qn[2], qn[3], qn[4], qn[5], 1.0, &s.zee[0]) != GM_OK)
        return fail("unsupported Zeeman pattern");
    /* PSF (gmilos.cu:242-301) */
    s.use_psf = cfg.convolve_with_psf;
    if (s.use_psf) {
        if (cfg.fwhm > 0) {
            gm_psf_gaussian(cfg.fwhm, lam.data(), n, s.psf);
        } else {
            const int rows = gmio_count_lines(cfg.psf_file);
            if (rows <= 0) return fail("empty or unreadable PSF file ", cfg.psf_file);
            std::vector<double> dl(rows), val(rows);
            const int got = gmio_read_psf(cfg.psf_file, dl.data(), val.data(), rows);
            if (got < 2 || gm_psf_from_table(dl.data(), val.data(), got, off.data(), n, step, s.psf) != GM_OK)
                return fail("the PSF table does not cover the grid: ", cfg.psf_file);
        }
    }
    for (int i = 0; i < 4; ++i) s.weight[i] = cfg.weight[i];
    s.noise = cfg.noise;
    s.ilambda = cfg.initial_diagonal_element;
    s.toplim = cfg.toplim;
    s.miter = cfg.number_of_cycles > 0 ? cfg.number_of_cycles : 1;
    s.mu = cfg.mu;
    s.logclambda = cfg.logclambda;
    for (int i = 0; i < 11; ++i) s.fix[i] = cfg.fix[i];
    for (int i = 0; i < 11; ++i) s.init_model[i] = init[i];
    s.precision = precision;
    s.solver = solver;
    s.rf_reference_rules = exact_rf ? 0 : 1;
    /* stray light: read for the synthesis mode when the file exists
     * (gmilos.cu:434-461) and, for inversions, only with a free filling
     * factor (gmilos.cu:688-716, 916-944) */
    bool use_stray_map = false;
    if (cfg.stray_light_file[0] && (cfg.number_of_cycles == 0 || cfg.fix[10])) {
        const char *se = gmio_file_ext(cfg.stray_light_file);
        if (!std::strcmp(se, ".per")) {
            std::vector<float> sl(4 * n);
            const int rc = gmio_read_stray_per(cfg.stray_light_file, n, off.data(), sl.data());
            if (rc == -3)
                return fail("the observed profiles and the stray light profiles do not have the "
                            "same wavelengths: ", cfg.stray_light_file);
            if (rc == 0) {
                s.use_slight = 1;
                for (int k = 0; k < 4 * n; ++k) s.slight[k] = sl[k];
            }
        } else if (!std::strcmp(se, ".fits")) {
            /* readFitsStrayLightFile (utilsFits.c:1578-1897): a 2-D float
             * image is one profile (wavelength the longer axis), a 4-D one a
             * per-pixel map laid out like the observed cube */
            int32_t naxis = 0;
            int64_t naxes[4] = {1, 1, 1, 1};
            if (gmio_fits_read_image(cfg.stray_light_file, &naxis, naxes, nullptr, 0) != 0)
                return fail("cannot read stray light FITS (float32): ", cfg.stray_light_file);
            if (naxis == 2) {
                const int64_t nl = naxes[0] > naxes[1] ? naxes[0] : naxes[1];
                const int64_t ns = naxes[0] > naxes[1] ? naxes[1] : naxes[0];
                if (nl != n || ns != 4) {
                    std::printf("The number of wavelengths is different in the stray light file: "
                                "%lld and the grid %d; stray light is not used.\n",
                                (long long)nl, n);
                } else {
                    std::vector<float> raw(4 * n);
                    gmio_fits_read_image(cfg.stray_light_file, &naxis, naxes, raw.data(), 4 * n);
                    s.use_slight = 1;
                    for (int st = 0; st < 4; ++st)
                        for (int k = 0; k < n; ++k)
                            s.slight[st * n + k] = (naxes[0] == nl) ? raw[st * n + k] : raw[k * 4 + st];
                }
            } else if (naxis == 4) {
                if (cfg.number_of_cycles == 0 ||
                    std::strcmp(gmio_file_ext(cfg.observed_profiles), ".fits")) {
                    std::printf("A 4-D stray light map needs a FITS cube to invert; stray light "
                                "is not used.\n");
                } else {
                    use_stray_map = true;
                    s.use_slight = 1;          /* the map replaces the profile */
                }
            } else {
                return fail("stray light FITS must have 2 or 4 axes: ", cfg.stray_light_file);
            }
        }
    }
    if (cfg.number_of_cycles > 0 && cfg.fix[10] && !s.use_slight)
        return fail("a free filling factor needs a stray light profile (.per): ",
                    cfg.stray_light_file);

    gm_ctx *ctx = nullptr;
    gm_status st = gm_ctx_create(&s, &ctx);
    if (st != GM_OK) return fail("engine set-up failed: ", gm_status_string(st));
    const std::string base = base_of(cfg);
    const char *ext = gmio_file_ext(cfg.observed_profiles);
    const bool is_per = !std::strcmp(ext, ".per"), is_fits = !std::strcmp(ext, ".fits");

    /* ---- synthesis ------------------------------------------------------- */
    if (cfg.number_of_cycles == 0) {
        Dev dm(sizeof(double) * 11), ds(sizeof(float) * 4 * n);
        std::vector<float> sp(4 * n);
        if (!dm.p || !ds.p) return fail("device allocation failed");
        (void)hipMemcpy(dm.p, init, sizeof init, hipMemcpyHostToDevice);
        st = gm_synth(ctx, (double *)dm.p, 1, (float *)ds.p, nullptr);
        (void)hipMemcpy(sp.data(), ds.p, sizeof(float) * 4 * n, hipMemcpyDeviceToHost);
        if (st != GM_OK) return fail("synthesis failed: ", gm_status_string(st));
        const std::string out = base + ".per";
        std::vector<double> rel(n);
        for (int k = 0; k < n; ++k) rel[k] = (lam[k] - lambda0) * 1000;
        gmio_write_per(out.c_str(), index, n, rel.data(), sp.data());
        std::printf("SYNTHESIS DONE: %s\n", out.c_str());
        gm_ctx_destroy(ctx);
        return 0;
    }

    /* ---- observations ---------------------------------------------------- */
    std::vector<float> obs;
    int64_t npix = 0;
    gmio_cube *cube = nullptr;
    int32_t nrows = 1, ncols = 1;
    if (is_per) {
        obs.resize(4 * n);
        if (gmio_read_per(cfg.observed_profiles, n, obs.data()) != 0)
            return fail("cannot read profiles ", cfg.observed_profiles);
        npix = 1;
    } else if (is_fits) {
        if (gmio_fits_read_cube(cfg.observed_profiles, n, &cube) != 0)
            return fail("cannot read FITS cube (float32, 4 axes, Stokes = 4, matching grid): ",
                        cfg.observed_profiles);
        int32_t nl;
        gmio_cube_dims(cube, &nrows, &ncols, &nl);
        npix = (int64_t)nrows * ncols;
        obs.assign(gmio_cube_data(cube), gmio_cube_data(cube) + npix * 4 * n);
    } else {
        return fail("observed profiles must be .per or .fits: ", cfg.observed_profiles);
    }
    Dev dobs(sizeof(float) * obs.size()), dmod(sizeof(double) * 11 * npix);
    if (!dobs.p || !dmod.p) return fail("device allocation failed");
    (void)hipMemcpy(dobs.p, obs.data(), sizeof(float) * obs.size(), hipMemcpyHostToDevice);
    /* per-pixel stray light (mapStrayLight, gmilos.cu:1212-1226) */
    Dev dmap(use_stray_map ? sizeof(float) * obs.size() : 0);
    if (use_stray_map) {
        gmio_cube *mc = nullptr;
        int32_t mr = 0, mcn = 0, ml = 0;
        if (gmio_fits_read_cube(cfg.stray_light_file, n, &mc) != 0)
            return fail("cannot read the stray light map: ", cfg.stray_light_file);
        gmio_cube_dims(mc, &mr, &mcn, &ml);
        if (mr != nrows || mcn != ncols) {
            gmio_cube_free(mc);
            return fail("the stray light map and the observed cube differ in size: ",
                        cfg.stray_light_file);
        }
        (void)hipMemcpy(dmap.p, gmio_cube_data(mc), sizeof(float) * obs.size(),
                        hipMemcpyHostToDevice);
        gmio_cube_free(mc);
        st = gm_ctx_set_stray_map(ctx, (const float *)dmap.p, npix);
        if (st != GM_OK) return fail("stray light map: ", gm_status_string(st));
    }
    std::vector<double> models(11 * npix);

    /* ---- classical estimates (gmilos.cu:309-431) -------------------------- */
    if (cfg.number_of_cycles < 0) {
        st = gm_classical_estimates(ctx, (float *)dobs.p, npix, 0, (double *)dmod.p, nullptr);
        (void)hipMemcpy(models.data(), dmod.p, sizeof(double) * models.size(), hipMemcpyDeviceToHost);
        if (st != GM_OK) return fail("estimates failed: ", gm_status_string(st));
        if (is_per) {
            /* the reference starts this mode from a zeroed model */
            for (int f : {0, 9, 3, 4, 10, 8}) models[f] = 0.0;
            const std::string out = base + "_model_ce.mod";
            gmio_write_mod(out.c_str(), models.data(), 0, 0.0);
            std::printf("CLASSICAL ESTIMATES SAVED IN FILE: %s\n", out.c_str());
        } else {
            std::vector<float> c2(npix, 0.f);
            std::vector<int32_t> it(npix, 0);
            const std::string out = base + "_model_ce.fits";
            gmio_fits_write_models(out.c_str(), nrows, ncols, models.data(), c2.data(), it.data(),
                                   cfg.save_chisqr);
            std::printf("CLASSICAL ESTIMATES SAVED IN FILE: %s\n", out.c_str());
        }
        gmio_cube_free(cube);
        gm_ctx_destroy(ctx);
        return 0;
    }

    /* ---- inversion ------------------------------------------------------- */
    Dev dc2(sizeof(float) * npix), dit(sizeof(int32_t) * npix), dsyn(sizeof(float) * obs.size());
    if (!dc2.p || !dit.p || !dsyn.p) return fail("device allocation failed");
    gm_out out = {(double *)dmod.p, (float *)dc2.p, (int32_t *)dit.p, (float *)dsyn.p, nullptr};
    auto t0 = std::chrono::steady_clock::now();
    st = gm_invert(ctx, (float *)dobs.p, npix, nullptr, &out, nullptr);
    (void)hipDeviceSynchronize();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (st != GM_OK) return fail("inversion failed: ", gm_status_string(st));
    std::vector<float> c2(npix), syn(obs.size());
    std::vector<int32_t> it(npix);
    (void)hipMemcpy(models.data(), dmod.p, sizeof(double) * models.size(), hipMemcpyDeviceToHost);
    (void)hipMemcpy(c2.data(), dc2.p, sizeof(float) * npix, hipMemcpyDeviceToHost);
    (void)hipMemcpy(it.data(), dit.p, sizeof(int32_t) * npix, hipMemcpyDeviceToHost);
    (void)hipMemcpy(syn.data(), dsyn.p, sizeof(float) * syn.size(), hipMemcpyDeviceToHost);
    std::printf("FINISH EXECUTION OF INVERSION: %f seconds to execute (%lld pixels)\n", secs,
                (long long)npix);
    if (is_per) {
        const std::string mo = base + "_model.mod", so = base + "_stokes.per";
        gmio_write_mod(mo.c_str(), models.data(), it[0], c2[0]);
        std::vector<double> rel(n);
        for (int k = 0; k < n; ++k) rel[k] = (lam[k] - lambda0) * 1000;
        gmio_write_per(so.c_str(), index, n, rel.data(), syn.data());
        std::printf("INVERTED MODEL SAVED IN FILE: %s\n", mo.c_str());
    } else {
        const std::string mo = base + "_mod.fits";
        if (gmio_fits_write_models(mo.c_str(), nrows, ncols, models.data(), c2.data(), it.data(),
                                   cfg.save_chisqr) != 0)
            return fail("cannot write ", mo.c_str());
        std::printf("INVERTED MODEL SAVED IN FILE: %s\n", mo.c_str());
        if (cfg.save_synthesis_adjusted) {
            const std::string so = base + "_stokes.fits";
            if (gmio_fits_write_profiles(so.c_str(), cube, syn.data()) != 0)
                return fail("cannot write ", so.c_str());
        }
    }
    gmio_cube_free(cube);
    gm_ctx_destroy(ctx);
    return 0;
}
