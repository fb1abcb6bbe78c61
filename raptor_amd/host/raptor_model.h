/*
 * raptor_model.h -- RAPTOR's model hooks and pixel loop on top of the
 * MI355X library.  Same names and argument meaning as the reference's
 * (functions.h:20-25, 70), with its globals replaced by explicit arguments:
 * the primitive array `real ****p` is the rt_ctx's copy in HBM.
 */
#ifndef RAPTOR_MODEL_H
#define RAPTOR_MODEL_H

#include "raptor_io.h"
#include "raptor_rt.h"

/* init_model (raptor_harm_model.c:59-66): units and dump upload. */
int init_model(rt_ctx *ctx, const rio_model *m, const rio_dump *d);

/* get_fluid_params (raptor_harm_model.c:190-320): returns 1, or 0 when
 * b^2/rho > 1 (Ne, Thetae, B = 1e-20, as the reference sets them).  On a
 * library error it also returns 0 with Ne, Thetae, B = 1e-20, so that a
 * caller testing it with if() skips the sample, and sets *IN_VOLUME = 0 (the
 * reference always sets 1); rt_last_error(ctx) holds the message.
 * gcov/gcon may be NULL. */
int get_fluid_params(rt_ctx *ctx, const rt_grid *g, const double X[4], double *Ne,
                     double *Thetae, double *B, double *beta, double *Bern, double Bcon[4],
                     double Ucon[4], int *IN_VOLUME, double gcov[4][4], double gcon[4][4]);

/* calculate_image (core.c:74-152) for one inclination: fills
 * intensityfield[y*W + x][f] and adds the pixel sums to energy_spectrum[f]. */
int calculate_image(rt_ctx *ctx, const rio_model *m, double incl, double **intensityfield,
                    double *energy_spectrum, const double *frequencies);

#endif
