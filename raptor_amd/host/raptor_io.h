/*
 * raptor_io.h -- host side of the RAPTOR command-line driver: model.in, the
 * HARM dump, and the image/spectrum files, in the reference's formats.
 */
#ifndef RAPTOR_IO_H
#define RAPTOR_IO_H

#include <stdio.h>

#include "raptor_rt.h"

#define RIO_MAX_FREQ RT_MAX_FREQ
#define RIO_MAX_CAM RT_MAX_CAMERAS

/* Everything read_model (core.c:4-72) reads into globals. */
typedef struct {
    double MBH, M_UNIT;
    int ABSORPTION;               /* read, unused by the reference too */
    int IMG_WIDTH, IMG_HEIGHT;
    double CAM_SIZE_X, CAM_SIZE_Y;
    int FREQS_PER_DEC;
    double FREQ_MIN, FREQ_MAX;    /* FREQ_MAX is read and unused (core.c:33) */
    double STEPSIZE;
    char GRMHD_FILE[256];
    double INCLINATION, R_HIGH, R_LOW, TIME_INIT;
    /* extensions: frequency and inclination lists */
    int n_freq;
    double freqs[RIO_MAX_FREQ];
    int n_cam;
    double incl[RIO_MAX_CAM];
    int device;
    int integrator;               /* RT_INT_* (int_method, parameters.h:186-189) */
    int emission;                 /* RT_EMIS_* (j_nu.c emission models) */
} rio_model;

/* A primitive set in the reference's index order p[v][i][j][k]. */
typedef struct {
    rt_grid grid;
    double gam, Rin, Rout, t;
    float *prims;                 /* 8 * n1 * n2 * n3 */
    /* accretion diagnostics of an ASCII dump (raptor_harm_model.c:171-180),
     * code units: rest-mass flux through the inner 21 radial zones and the
     * advected energy flux through zones 20..39, per 21 zones */
    int have_fluxes;
    double dMact, Ladv;
} rio_dump;

/* model.in (core.c:22-34): 11 records "NAME (unit) value" in fixed order.
 * Returns 0 or -1 with a message in err. */
int rio_read_model_in(const char *path, rio_model *m, char *err, size_t errlen);

/* argv overrides (core.c:37-42) and the optional flags
 *   --freqs f1,f2,...   explicit frequency list (Hz)
 *   --nfreq N           N frequencies FREQ_MIN * 10^(f / FREQS_PER_DEC)
 *   --incl i1,i2,...    several inclinations in one launch
 *   --device D
 *   --integrator rk2|rk4|verlet   the stepper (int_method); rk2 by default
 *   --emission thsynch|thsynchav|kappa   the emission model; thsynch (the
 *                       one emission_coeff calls) by default
 * argv: prog model.in dump M_UNIT INCLINATION R_HIGH R_LOW TIME_INIT [flags] */
int rio_parse_args(int argc, char **argv, rio_model *m, char *err, size_t errlen);

/* Reads a 2D HARM ASCII dump (raptor_harm_model.c:90-176) or a binary
 * primitive file (magic "RPTRPRIM", any N3); the format is detected from the
 * first bytes. */
int rio_read_dump(const char *path, rio_dump *d, char *err, size_t errlen);
/* The reference's line after reading the dump (raptor_harm_model.c:181-182):
 * dMact, the accretion rate in Msun/year (dMact M_unit / T_unit with
 * T_unit = G MBH / c^3) and Ladv, then "Done reading data".  Prints nothing
 * but the last line for dumps without the fluxes. */
void rio_print_fluxes(FILE *fp, const rio_dump *d, double M_unit, double MBH);
void rio_free_dump(rio_dump *d);

/* JANSKY_FACTOR (utilities.c:32-42). */
double rio_jansky_factor(const rio_model *m);

/* output_files (core.c:154-200) for one inclination: image files
 * output/img_data_<t>_<freq>_<incl>.dat and output/spectrum_<t>_<incl>.dat.
 * image is [W*H][n_freq] (pixel index y*W + x). */
int rio_write_outputs(const rio_model *m, double incl, const double *image,
                      const double *spectrum, const char *outdir, char *err, size_t errlen);

#endif
