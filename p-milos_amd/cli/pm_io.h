/*
 * pm_io.h -- the file surface of the command line tool (host code only):
 * the control files (.mtrol, .minit), line list (LINES), wavelength grid
 * (.grid or a 2-D wavelength FITS), model files (.mod), PSF files (.psf),
 * profile files (.per) and FITS cubes, written from the formats the
 * reference reads and writes.
 *
 * Reference interfaces replaced (P-MILOS, src/):
 *   pmio_read_mtrol     readTrolFile + loadInitialValues, readConfig.c:371-1023
 *   pmio_read_minit     readInitFile, readConfig.c:1032-1227
 *   pmio_read_lines     readFileCuanticLines, readConfig.c:24-129
 *   pmio_read_grid      readMallaGrid, readConfig.c:295-335
 *   pmio_read_model     readInitialModel + checkInitialModel, readConfig.c:136-293
 *   pmio_read_psf_file  readPSFFile, readConfig.c:341-368
 *   pmio_interp_psf     interpolationLinearPSF, milosUtils.c:684-738
 *   pmio_read_per       the .per reader of milos.c:654-688 / utilsFits.c:835-877
 *   pmio_fits_*         readFitsSpectroImage (utilsFits.c:25-364),
 *                       readFitsStrayLightFile (:881-1200), readFitsMaskFile
 *                       (:1541-1677), readFitsLambdaToArray (:749-830),
 *                       writeFitsImageModels (:1706-1823),
 *                       writeFitsImageProfiles (:1952-2088)
 * Every function returns 0 on success and a negative code (or a count) as
 * documented; none of them exits the process.
 */
#ifndef PM_IO_H
#define PM_IO_H

#ifdef __cplusplus
extern "C" {
#endif

#define PMIO_PATH 4096

/* The fields of ConfigControl (defines.h:223-277) the tools use. */
typedef struct {
    /* .mtrol */
    int cycles;                      /* NumberOfCycles: 0 synthesis, <0 estimates, >0 max iterations */
    char observed[PMIO_PATH];        /* ObservedProfiles */
    char straylight[PMIO_PATH];      /* StrayLightFile */
    char psf[PMIO_PATH];             /* PSF file, or the FWHM text */
    double fwhm;                     /* > 0: Gaussian PSF of this FWHM (mA) */
    int use_psf;                     /* ConvolveWithPSF */
    int psf_is_file;                 /* a .psf file */
    char wavelength[PMIO_PATH];      /* wavelength FITS, when not a grid */
    char grid[PMIO_PATH];            /* MallaGrid */
    int use_grid;
    char lines[PMIO_PATH];           /* AtomicParametersFile */
    char guess[PMIO_PATH];           /* InitialGuessModel */
    double weight[4];                /* WeightForStokes */
    int fix[11];                     /* fix[] in Init_Model order */
    int invert_filling;              /* InvertFillingFactor (read, unused) */
    double mu;
    int sn;                          /* EstimatedSNForI */
    double noise;                    /* 1/SN^2 */
    double ilambda;                  /* InitialDiagonalElement, read through a float */
    int use_fft;
    int logclambda;
    /* .minit */
    char control[PMIO_PATH];
    int subx1, subx2, suby1, suby2;
    char prefix[PMIO_PATH];          /* outputPrefix */
    char mask[PMIO_PATH];            /* MaskFile */
    int t1, t2;
    int invert_directory;            /* t1 = '*' */
    int loop_inversion;              /* t2 = '*' */
    int save_synthesis;              /* SaveSynthesisAdjusted */
} pmio_control;

/* Defaults of loadInitialValues, then the 42 lines of a .mtrol in order
 * (blank values keep the default, as sscanf leaves them).  Returns 0, -1 if
 * the file cannot be read, -2 if it is shorter than 42 lines, -3 for a
 * wavelength file or PSF file of an unknown kind. */
int pmio_read_mtrol(const char *path, pmio_control *out);

/* The 12 lines of a .minit; reads the .mtrol it names first.  Returns 0 or
 * a negative code (-4: a subfield with some but not all bounds given). */
int pmio_read_minit(const char *path, pmio_control *out);

/* Line `index` of a LINES file: dat = {1, S_l, L_l, J_l, S_u, L_u, J_u}
 * (spin from the multiplicity, (2S+1-1)/2); returns lambda0 (A), 0 if the
 * line is not there. */
double pmio_read_lines(const char *path, int index, double dat[7]);

/* The grid line after the `------` header line: returns the line index
 * (> 0), 0 if there is none. */
int pmio_read_grid(const char *path, double *initial_mA, double *step_mA, double *final_mA);

/* A .mod file into model[11] (Init_Model order; fields missing from the
 * file keep their values).  Returns 0, -1 if unreadable, or 1 + the index of
 * the first parameter outside the reference's accepted range. */
int pmio_read_model(const char *path, double model[11]);

/* A .psf file (two columns: offset in mA, value): returns the number of
 * samples read into dl/psf (at most cap), -1 if unreadable. */
int pmio_read_psf_file(const char *path, double *dl, double *psf, int cap);

/* Linear interpolation of the PSF samples at offsets_mA[i] - shift_mA, zero
 * outside the sampled range, normalised to unit sum. */
int pmio_interp_psf(const double *dl, const double *psf, int npsf, const double *offsets_mA,
                    int n, double shift_mA, double *G);

/* A .per profile file: L lines of "index offset I Q U V" into spec[4][L]
 * (float, Stokes-major).  offsets_mA (nullable) receives the offsets.
 * Returns the number of lines, -1 if unreadable. */
int pmio_read_per(const char *path, int L, float *spec, double *offsets_mA);

/* Writes "index\toffset\tI\tQ\tU\tV" lines (%d\t%f\t%e...) for spec[4][L]. */
int pmio_write_per(const char *path, int index, const double *offsets_mA, const double *spec, int L);

/* The model text file of the .per modes (milos.c:787-808). */
int pmio_write_model_text(const char *path, const double model[11], int iters, double chi2,
                          int estimates_layout);

/* ---- FITS (primary HDU images) ---- */

typedef struct {
    int naxis;
    long naxes[4];
    int pos_lambda, pos_stokes, pos_row, pos_col;   /* axis of each coordinate */
    int rows, cols, nlambda;
    int ncards;                                     /* header cards kept for copies */
    char (*cards)[81];
} pmio_cube_layout;

/* Reads a float32 4-D Stokes cube into data[npix][4][L] (malloc'd; pixel
 * index = col * rows + row) with the axes found from CTYPE1-4 (WAVE-GRI,
 * STOKES, HPLN-TAN, HPLT-TAN) or, without them, from the axis sizes as the
 * reference guesses them.  Any axis order is read by its strides.  A
 * subfield (x = rows, y = cols, 1-based inclusive; all 0 = whole image) is
 * cut out.  Returns 0, -1 unreadable, -2 not float32, -3 not 4-D with 4
 * Stokes, -4 wavelength count != L, -5 bad subfield. */
int pmio_fits_read_cube(const char *path, int L, int subx1, int subx2, int suby1, int suby2,
                        float **data, pmio_cube_layout *layout);
void pmio_free(void *p);
void pmio_free_layout(pmio_cube_layout *layout);

/* Writes the fit cube with the layout (axis order, header cards) of the
 * cube it was read from; data[npix][4][L]. */
int pmio_fits_write_cube(const char *path, const pmio_cube_layout *layout, const float *data);

/* The 13-plane model cube: NAXIS1 = rows, NAXIS2 = cols, NAXIS3 = 13 planes
 * (eta0 B vlos dopp aa gm az S0 S1 mac alfa iterations chi2), each plane in
 * pixel order; models [npix][11]; chi2/iters nullable (zeros). */
int pmio_fits_write_models(const char *path, int rows, int cols, const double *models,
                           const float *chi2, const int *iters);

/* A 2-D image of any BITPIX as doubles (malloc'd), naxes[0] fastest.
 * Returns 0 or a negative code. */
int pmio_fits_read_2d(const char *path, double **data, long *n1, long *n2);

/* A stray-light FITS: 2-D -> one profile [4][L] (*per_pixel = 0); 4-D ->
 * one profile per pixel [npix][4][L] (*per_pixel = 1), read like a Stokes
 * cube. */
int pmio_fits_read_straylight(const char *path, int L, int subx1, int subx2, int suby1,
                              int suby2, float **data, int *per_pixel, long *npix);

/* A 2-D wavelength FITS: NAXIS1 = L; row 1 the line index, row 2 the
 * wavelengths (A).  lambda (malloc'd) [L]. */
int pmio_fits_read_lambda(const char *path, double **lambda, int *L, int *line_index);

#ifdef __cplusplus
}
#endif
#endif
