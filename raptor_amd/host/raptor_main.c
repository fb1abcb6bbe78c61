/*
 * raptor_main.c -- the RAPTOR command line (main.c:95-167) on one MI355X.
 *
 *   raptor model.in dumpfile M_UNIT INCLINATION R_HIGH R_LOW TIME_INIT
 *          [--freqs f1,f2,..] [--nfreq N] [--incl i1,i2,..] [--device D]
 *
 * Reads model.in and the dump like the reference, renders every inclination
 * in one launch of the library, and writes output/img_data_*.dat and
 * output/spectrum_*.dat in the reference's formats.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "raptor_io.h"
#include "raptor_model.h"
#include "raptor_rt.h"

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char **argv)
{
    char err[512];
    rio_model m;
    memset(&m, 0, sizeof m);
    fprintf(stderr, "WELCOME TO RAPTOR (MI355X)\n\n");
    if (argc < 2) {
        rio_parse_args(argc, argv, &m, err, sizeof err);
        fprintf(stderr, "%s\n", err);
        return 2;
    }
    if (rio_read_model_in(argv[1], &m, err, sizeof err) || rio_parse_args(argc, argv, &m, err, sizeof err)) {
        fprintf(stderr, "%s\n", err);
        return 2;
    }
    fprintf(stderr, "Model parameters:\nMBH \t\t= %g \nM_UNIT \t\t= %g \nR_HIGH \t\t= %g \n"
                    "R_LOW \t\t= %g \nINCLINATION \t= %g \n",
            m.MBH, m.M_UNIT, m.R_HIGH, m.R_LOW, m.INCLINATION);
    fprintf(stderr, "Observer parameters:\nIMG_WIDTH \t= %d \nIMG_HEIGHT \t= %d \nCAM_SIZE_X \t= %g \n"
                    "CAM_SIZE_Y \t= %g \nNUM OF FREQS \t= %d \nSTEPSIZE \t= %g \n",
            m.IMG_WIDTH, m.IMG_HEIGHT, m.CAM_SIZE_X, m.CAM_SIZE_Y, m.n_freq, m.STEPSIZE);

    rio_dump d;
    double t0 = now_s();
    if (rio_read_dump(m.GRMHD_FILE, &d, err, sizeof err)) {
        fprintf(stderr, "%s\n", err);
        return 1;
    }
    fprintf(stderr, "Read %s: %d x %d x %d cells (%.2f s)\n", m.GRMHD_FILE, d.grid.n1,
            d.grid.n2, d.grid.n3, now_s() - t0);
    rio_print_fluxes(stderr, &d, m.M_UNIT, m.MBH);

    rt_ctx *ctx = rt_create(m.device);
    if (!ctx) {
        fprintf(stderr, "rt_create: %s\n", rt_last_error(NULL));
        rio_free_dump(&d);
        return 1;
    }
    int rc = init_model(ctx, &m, &d);
    if (rc != RT_OK) {
        fprintf(stderr, "init_model: %s\n", rt_last_error(ctx));
        rt_destroy(ctx);
        rio_free_dump(&d);
        return 1;
    }
    size_t npix = (size_t)m.IMG_WIDTH * m.IMG_HEIGHT;
    size_t per_cam = npix * m.n_freq;
    double *image = malloc(per_cam * m.n_cam * sizeof(double));
    double *spec = malloc((size_t)m.n_freq * m.n_cam * sizeof(double));
    if (!image || !spec) {
        fprintf(stderr, "out of host memory\n");
        return 1;
    }
    double t1 = now_s();
    rc = rt_render(ctx, m.incl, m.n_cam, m.freqs, m.n_freq, image, NULL);
    double t2 = now_s();
    if (rc != RT_OK) {
        fprintf(stderr, "rt_render: %s\n", rt_last_error(ctx));
        return 1;
    }
    /* core.c:146-150: each image summed pixel by pixel in the reference's
     * order, so that the spectrum file has the reference's digits */
    for (int k = 0; k < m.n_cam; k++)
        for (int f = 0; f < m.n_freq; f++) {
            double s = 0.;
            for (size_t i = 0; i < npix; i++)
                s += image[k * per_cam + i * m.n_freq + f];
            spec[k * m.n_freq + f] = s;
        }
    fprintf(stderr, "\nFinished ray tracing: %zu rays in %.3f s (kernel %.3f ms), "
                    "speed: %.4g [geodesics/sec]\n",
            npix * m.n_cam, t2 - t1, rt_last_kernel_ms(ctx), npix * m.n_cam / (t2 - t1));
    for (int k = 0; k < m.n_cam; k++) {
        if (rio_write_outputs(&m, m.incl[k], image + k * per_cam, spec + k * m.n_freq, "output",
                              err, sizeof err)) {
            fprintf(stderr, "%s\n", err);
            return 1;
        }
    }
    free(image);
    free(spec);
    rt_destroy(ctx);
    rio_free_dump(&d);
    return 0;
}
