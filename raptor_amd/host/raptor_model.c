/*
 * raptor_model.c -- the reference's model hooks and pixel loop, re-seated on
 * the MI355X library (include/raptor_rt.h).
 *
 *   init_model        raptor_harm_model.c:59-66  (units + dump -> HBM)
 *   get_fluid_params  raptor_harm_model.c:190-320 (evaluated by the device sampler)
 *   calculate_image   core.c:74-152              (one launch for all pixels)
 *
 * The reference passes its state through globals; here it travels in an
 * rt_ctx plus the rio_model read from model.in and argv.
 */
#include "raptor_model.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void fill_params(const rio_model *m, rt_params *p)
{
    memset(p, 0, sizeof *p);
    p->MBH = m->MBH;
    p->M_unit = m->M_UNIT;
    p->R_high = m->R_HIGH;
    p->R_low = m->R_LOW;
    p->width = m->IMG_WIDTH;
    p->height = m->IMG_HEIGHT;
    p->cam_size_x = m->CAM_SIZE_X;
    p->cam_size_y = m->CAM_SIZE_Y;
    p->stepsize = m->STEPSIZE;
    p->max_steps = 100000; /* parameters.h:182 */
    p->integrator = m->integrator;
    p->emission = m->emission;
}

int init_model(rt_ctx *ctx, const rio_model *m, const rio_dump *d)
{
    rt_params p;
    fill_params(m, &p);
    int rc = rt_set_params(ctx, &p);
    if (rc != RT_OK)
        return rc;
    /* grmhd.c:17-33 set_units, printed like the reference */
    double L = 6.6742e-8 * m->MBH / (2.99792458e10 * 2.99792458e10);
    double rho = m->M_UNIT / pow(L, 3);
    fprintf(stderr, "\nUNITS\nL,T,M: %g %g %g\nrho,u,B: %g %g %g\n", L, L / 2.99792458e10,
            m->M_UNIT, rho, rho * 2.99792458e10 * 2.99792458e10,
            2.99792458e10 * sqrt(4. * M_PI * rho));
    fprintf(stderr, "Sim range x1, x2:  %g %g, %g %g\n", d->grid.startx1,
            d->grid.startx1 + d->grid.n1 * d->grid.dx1, d->grid.startx2,
            d->grid.startx2 + d->grid.n2 * d->grid.dx2);
    return rt_load_prims(ctx, &d->grid, d->prims);
}

/* MKS metric (metric.c:94-138, 207-231) for the hook's gcov/gcon side output */
static void mks_metric(const rt_grid *g, const double X[4], double gcov[4][4], double gcon[4][4])
{
    double a = g->a, h = g->hslope;
    double r = exp(X[1]);
    double th = M_PI * X[2] + 0.5 * (1. - h) * sin(2. * M_PI * X[2]);
    double s = sin(th), s2 = s * s, c = cos(th);
    double rho2 = r * r + a * a * c * c;
    double hfac = M_PI + (1. - h) * M_PI * cos(2. * M_PI * X[2]);
    memset(gcov, 0, 16 * sizeof(double));
    memset(gcon, 0, 16 * sizeof(double));
    gcov[0][0] = -1. + 2. * r / rho2;
    gcov[0][1] = gcov[1][0] = 2. * r / rho2 * r;
    gcov[0][3] = gcov[3][0] = -2. * a * r * s2 / rho2;
    gcov[1][1] = (1. + 2. * r / rho2) * r * r;
    gcov[1][3] = gcov[3][1] = -a * s2 * (1. + 2. * r / rho2) * r;
    gcov[2][2] = rho2 * hfac * hfac;
    gcov[3][3] = s2 * (rho2 + a * a * s2 * (1. + 2. * r / rho2));
    double ir = 1. / rho2;
    gcon[0][0] = -1. - 2. * r * ir;
    gcon[0][1] = gcon[1][0] = 2. * ir;
    gcon[1][1] = ir * (r * (r - 2.) + a * a) / (r * r);
    gcon[1][3] = gcon[3][1] = a * ir / r;
    gcon[2][2] = ir / (hfac * hfac);
    gcon[3][3] = ir / s2;
}

int get_fluid_params(rt_ctx *ctx, const rt_grid *g, const double X[4], double *Ne,
                     double *Thetae, double *B, double *beta, double *Bern, double Bcon[4],
                     double Ucon[4], int *IN_VOLUME, double gcov[4][4], double gcon[4][4])
{
    double out[14];
    if (rt_fluid_params(ctx, X, 1, out) != RT_OK) {
        /* no emission, as for the sigma cut (raptor_harm_model.c:285-290),
         * so that callers testing the result with if() skip the sample; the
         * error shows as IN_VOLUME = 0 (the reference always sets 1) and in
         * rt_last_error(ctx) */
        *Ne = *Thetae = *B = 1e-20;
        *beta = *Bern = 0.;
        for (int i = 0; i < 4; i++)
            Ucon[i] = Bcon[i] = 0.;
        *IN_VOLUME = 0;
        return 0;
    }
    *Ne = out[0];
    *Thetae = out[1];
    *B = out[2];
    for (int i = 0; i < 4; i++) {
        Ucon[i] = out[3 + i];
        Bcon[i] = out[7 + i];
    }
    *beta = out[12];
    *Bern = out[13];
    *IN_VOLUME = 1;
    if (gcov && gcon)
        mks_metric(g, X, gcov, gcon);
    return out[11] != 0. ? 1 : 0;
}

int calculate_image(rt_ctx *ctx, const rio_model *m, double incl, double **intensityfield,
                    double *energy_spectrum, const double *frequencies)
{
    size_t npix = (size_t)m->IMG_WIDTH * m->IMG_HEIGHT;
    int nf = m->n_freq;
    double *img = malloc(npix * nf * sizeof(double));
    if (!img)
        return RT_E_NOMEM;
    int rc = rt_render(ctx, &incl, 1, frequencies, nf, img, NULL);
    if (rc == RT_OK) {
        /* core.c:136-150: the image, and the spectrum summed pixel by pixel
         * in the reference's order, so that the spectrum file has its digits */
        for (size_t i = 0; i < npix; i++)
            for (int f = 0; f < nf; f++) {
                intensityfield[i][f] = img[i * nf + f];
                energy_spectrum[f] += img[i * nf + f];
            }
    }
    free(img);
    return rc;
}
