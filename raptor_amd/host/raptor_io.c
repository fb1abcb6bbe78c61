/*
 * raptor_io.c -- model.in, HARM dump and output files in RAPTOR's formats.
 *
 * Formats kept byte-compatible with the reference:
 *   model.in records                core.c:22-34
 *   HARM 2D ASCII dump              raptor_harm_model.c:90-176
 *   img_data_*.dat, spectrum_*.dat  core.c:163-198, utilities.c:47-64
 */
#include "raptor_io.h"

#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

static int seterr(char *err, size_t n, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    if (err && n)
        vsnprintf(err, n, fmt, ap);
    va_end(ap);
    return -1;
}

int rio_read_model_in(const char *path, rio_model *m, char *err, size_t errlen)
{
    FILE *f = fopen(path, "r");
    if (!f)
        return seterr(err, errlen, "cannot read input file %s: %s", path, strerror(errno));
    char a[100], b[100];
    int ok = 1;
    ok &= fscanf(f, "%99s %99s %lf", a, b, &m->MBH) == 3;
    ok &= fscanf(f, "%99s %99s %lf", a, b, &m->M_UNIT) == 3;
    ok &= fscanf(f, "%99s %99s %d", a, b, &m->ABSORPTION) == 3;
    ok &= fscanf(f, "%99s %99s %d", a, b, &m->IMG_WIDTH) == 3;
    ok &= fscanf(f, "%99s %99s %d", a, b, &m->IMG_HEIGHT) == 3;
    ok &= fscanf(f, "%99s %99s %lf", a, b, &m->CAM_SIZE_X) == 3;
    ok &= fscanf(f, "%99s %99s %lf", a, b, &m->CAM_SIZE_Y) == 3;
    ok &= fscanf(f, "%99s %99s %d", a, b, &m->FREQS_PER_DEC) == 3;
    ok &= fscanf(f, "%99s %99s %lf", a, b, &m->FREQ_MIN) == 3;
    ok &= fscanf(f, "%99s %99s %lf", a, b, &m->FREQ_MAX) == 3;
    ok &= fscanf(f, "%99s %99s %lf", a, b, &m->STEPSIZE) == 3;
    fclose(f);
    if (!ok)
        return seterr(err, errlen, "%s: expected 11 records 'NAME (unit) value'", path);
    if (m->IMG_WIDTH < 1 || m->IMG_HEIGHT < 1 || !(m->STEPSIZE > 0.) || m->FREQS_PER_DEC == 0)
        return seterr(err, errlen, "%s: bad image size, STEPSIZE or FREQS_PER_DEC", path);
    return 0;
}

static int parse_list(const char *s, double *out, int maxn)
{
    int n = 0;
    const char *p = s;
    while (*p && n < maxn) {
        char *end;
        double v = strtod(p, &end);
        if (end == p)
            return -1;
        out[n++] = v;
        p = end;
        if (*p == ',')
            p++;
        else if (*p)
            return -1;
    }
    return *p ? -1 : n;
}

int rio_parse_args(int argc, char **argv, rio_model *m, char *err, size_t errlen)
{
    if (argc < 8)
        return seterr(err, errlen,
                      "usage: %s model.in dumpfile M_UNIT INCLINATION R_HIGH R_LOW TIME_INIT "
                      "[--freqs f1,f2,..] [--nfreq N] [--incl i1,i2,..] [--device D] "
                      "[--integrator rk2|rk4|verlet] [--emission thsynch|thsynchav|kappa]",
                      argc ? argv[0] : "raptor");
    if (strlen(argv[2]) >= sizeof m->GRMHD_FILE)
        return seterr(err, errlen, "dump file name too long");
    /* core.c:37-42 */
    strcpy(m->GRMHD_FILE, argv[2]);
    if (sscanf(argv[3], "%lf", &m->M_UNIT) != 1 || sscanf(argv[4], "%lf", &m->INCLINATION) != 1 ||
        sscanf(argv[5], "%lf", &m->R_HIGH) != 1 || sscanf(argv[6], "%lf", &m->R_LOW) != 1 ||
        sscanf(argv[7], "%lf", &m->TIME_INIT) != 1)
        return seterr(err, errlen, "M_UNIT, INCLINATION, R_HIGH, R_LOW, TIME_INIT must be numbers");
    /* main.c:128-132 with num_indices = 1 unless --nfreq / --freqs */
    m->n_freq = 1;
    m->freqs[0] = m->FREQ_MIN;
    m->n_cam = 1;
    m->incl[0] = m->INCLINATION;
    m->device = 0;
    m->integrator = RT_INT_RK2;
    m->emission = RT_EMIS_THSYNCH;
    for (int i = 8; i < argc; i++) {
        if (i + 1 >= argc)
            return seterr(err, errlen, "flag %s needs a value", argv[i]);
        const char *v = argv[++i];
        if (!strcmp(argv[i - 1], "--freqs")) {
            int n = parse_list(v, m->freqs, RIO_MAX_FREQ);
            if (n < 1)
                return seterr(err, errlen, "--freqs: 1..%d comma-separated numbers", RIO_MAX_FREQ);
            m->n_freq = n;
        } else if (!strcmp(argv[i - 1], "--nfreq")) {
            int n = atoi(v);
            if (n < 1 || n > RIO_MAX_FREQ)
                return seterr(err, errlen, "--nfreq must be in [1, %d]", RIO_MAX_FREQ);
            m->n_freq = n;
            for (int f = 0; f < n; f++)
                m->freqs[f] = m->FREQ_MIN * pow(10., (double)f / (double)m->FREQS_PER_DEC);
        } else if (!strcmp(argv[i - 1], "--incl")) {
            int n = parse_list(v, m->incl, RIO_MAX_CAM);
            if (n < 1)
                return seterr(err, errlen, "--incl: 1..%d comma-separated numbers", RIO_MAX_CAM);
            m->n_cam = n;
        } else if (!strcmp(argv[i - 1], "--device")) {
            m->device = atoi(v);
        } else if (!strcmp(argv[i - 1], "--integrator")) {
            if (!strcmp(v, "rk2"))
                m->integrator = RT_INT_RK2;
            else if (!strcmp(v, "rk4"))
                m->integrator = RT_INT_RK4;
            else if (!strcmp(v, "verlet"))
                m->integrator = RT_INT_VERLET;
            else
                return seterr(err, errlen, "--integrator must be rk2, rk4 or verlet");
        } else if (!strcmp(argv[i - 1], "--emission")) {
            if (!strcmp(v, "thsynch"))
                m->emission = RT_EMIS_THSYNCH;
            else if (!strcmp(v, "thsynchav"))
                m->emission = RT_EMIS_THSYNCHAV;
            else if (!strcmp(v, "kappa"))
                m->emission = RT_EMIS_KAPPA;
            else
                return seterr(err, errlen, "--emission must be thsynch, thsynchav or kappa");
        } else {
            return seterr(err, errlen, "unknown flag %s", argv[i - 1]);
        }
    }
    return 0;
}

/* ---- dumps ----------------------------------------------------------------- */

static int read_ascii(FILE *fp, rio_dump *d, char *err, size_t errlen)
{
    double hdr[26];
    for (int k = 0; k < 26; k++)
        if (fscanf(fp, "%lf", &hdr[k]) != 1)
            return seterr(err, errlen, "HARM header: expected 26 numbers");
    int n1 = (int)hdr[1], n2 = (int)hdr[2];
    if (n1 < 2 || n2 < 2 || (double)n1 != hdr[1] || (double)n2 != hdr[2])
        return seterr(err, errlen, "HARM header: bad n1 x n2 = %g x %g", hdr[1], hdr[2]);
    rt_grid *g = &d->grid;
    g->n1 = n1;
    g->n2 = n2;
    g->n3 = 1;
    g->startx1 = hdr[3];
    g->startx2 = hdr[4];
    g->startx3 = 0.;
    g->dx1 = hdr[5];
    g->dx2 = hdr[6];
    g->dx3 = 2. * M_PI;
    g->a = hdr[9];
    d->gam = hdr[10];
    d->Rin = hdr[22];
    d->Rout = hdr[23];
    g->hslope = hdr[24];
    g->R0 = hdr[25];
    d->t = hdr[0];
    size_t ncell = (size_t)n1 * n2;
    d->prims = malloc(8 * ncell * sizeof(float));
    if (!d->prims)
        return seterr(err, errlen, "out of memory for %zu cells", ncell);
    double rec[34];
    double dMact = 0., Ladv = 0.;
    for (size_t k = 0; k < ncell; k++) {
        for (int q = 0; q < 34; q++)
            if (fscanf(fp, "%lf", &rec[q]) != 1) {
                free(d->prims);
                d->prims = NULL;
                return seterr(err, errlen, "HARM dump: record %zu of %zu incomplete", k, ncell);
            }
        /* record k = i*n2 + j: x1 x2 r h, the primitives (4..11), divb,
         * U^mu (13..16), U_mu (17..20), B^mu, B_mu, 4 bounds, gdet (33) */
        for (int v = 0; v < 8; v++)
            d->prims[v * ncell + k] = (float)rec[4 + v];
        const size_t i = k / (size_t)n2;
        const double gdet = rec[33], ucon1 = rec[14], ucov0 = rec[17];
        if (i <= 20)
            dMact += gdet * rec[4] * ucon1;
        if (i >= 20 && i < 40)
            Ladv += gdet * rec[5] * ucon1 * ucov0;
    }
    /* over phi (dx3 = 2 pi for the 2D dump) and theta, per 21 zones */
    d->dMact = dMact * (2. * M_PI) * g->dx2 / 21.;
    d->Ladv = Ladv * (2. * M_PI) * g->dx2 / 21.;
    d->have_fluxes = 1;
    return 0;
}

void rio_print_fluxes(FILE *fp, const rio_dump *d, double M_unit, double MBH)
{
    if (d->have_fluxes) {
        const double c = 2.99792458e10, G = 6.6742e-8, msun = 1.989e33, year = 31536000.;
        const double T_unit = G * MBH / (c * c * c);
        fprintf(fp, "dMact: %g Macc: %e [Msun/year], Ladv: %g\n", d->dMact,
                d->dMact * M_unit / T_unit / (msun / year), d->Ladv);
    }
    fprintf(fp, "Done reading data\n\n");
}

static int read_bin(FILE *fp, rio_dump *d, char *err, size_t errlen)
{
    char magic[8];
    int32_t iv[5];
    double dv[12];
    if (fread(magic, 1, 8, fp) != 8 || fread(iv, sizeof(int32_t), 5, fp) != 5 ||
        fread(dv, sizeof(double), 12, fp) != 12)
        return seterr(err, errlen, "primitive file: short header");
    if (memcmp(magic, "RPTRPRIM", 8) != 0 || iv[0] != 1)
        return seterr(err, errlen, "primitive file: bad magic or version");
    rt_grid *g = &d->grid;
    g->n1 = iv[1];
    g->n2 = iv[2];
    g->n3 = iv[3];
    if (g->n1 < 2 || g->n2 < 2 || g->n3 < 1)
        return seterr(err, errlen, "primitive file: bad grid %d x %d x %d", g->n1, g->n2, g->n3);
    g->startx1 = dv[0];
    g->startx2 = dv[1];
    g->startx3 = dv[2];
    g->dx1 = dv[3];
    g->dx2 = dv[4];
    g->dx3 = dv[5];
    g->a = dv[6];
    d->gam = dv[7];
    g->hslope = dv[8];
    g->R0 = dv[9];
    d->Rin = dv[10];
    d->Rout = dv[11];
    d->t = 0.;
    size_t n = 8 * (size_t)g->n1 * g->n2 * g->n3;
    d->prims = malloc(n * sizeof(float));
    if (!d->prims)
        return seterr(err, errlen, "out of memory for %zu floats", n);
    if (fread(d->prims, sizeof(float), n, fp) != n) {
        free(d->prims);
        d->prims = NULL;
        return seterr(err, errlen, "primitive file truncated");
    }
    return 0;
}

int rio_read_dump(const char *path, rio_dump *d, char *err, size_t errlen)
{
    memset(d, 0, sizeof *d);
    FILE *fp = fopen(path, "rb");
    if (!fp)
        return seterr(err, errlen, "can't open sim data file %s: %s", path, strerror(errno));
    char head[8] = {0};
    size_t got = fread(head, 1, 8, fp);
    rewind(fp);
    int rc = (got == 8 && memcmp(head, "RPTRPRIM", 8) == 0) ? read_bin(fp, d, err, errlen)
                                                             : read_ascii(fp, d, err, errlen);
    fclose(fp);
    return rc;
}

void rio_free_dump(rio_dump *d)
{
    free(d->prims);
    d->prims = NULL;
}

/* ---- outputs ----------------------------------------------------------------- */

double rio_jansky_factor(const rio_model *m)
{
    /* utilities.c:21, 32-42 */
    double Rg = 6.6742e-8 * m->MBH / 2.99792458e10 / 2.99792458e10;
    double dx = m->CAM_SIZE_X * Rg / 2.6228263e22;
    double dy = m->CAM_SIZE_Y * Rg / 2.6228263e22;
    return 1.e23 * (dx / (double)m->IMG_WIDTH) * (dy / (double)m->IMG_HEIGHT);
}

int rio_write_outputs(const rio_model *m, double incl, const double *image,
                      const double *spectrum, const char *outdir, char *err, size_t errlen)
{
    struct stat st;
    if (stat(outdir, &st) == -1 && mkdir(outdir, 0700) != 0)
        return seterr(err, errlen, "cannot create %s: %s", outdir, strerror(errno));
    double jf = rio_jansky_factor(m);
    char name[1024];
    snprintf(name, sizeof name, "%s/spectrum_%d_%.02lf.dat", outdir, (int)m->TIME_INIT, incl);
    FILE *spec = fopen(name, "w");
    if (!spec)
        return seterr(err, errlen, "cannot write %s", name);
    const int W = m->IMG_WIDTH, H = m->IMG_HEIGHT, nf = m->n_freq;
    for (int f = 0; f < nf; f++) {
        snprintf(name, sizeof name, "%s/img_data_%d_%e_%.02lf.dat", outdir, (int)m->TIME_INIT,
                 m->freqs[f], incl);
        FILE *img = fopen(name, "w");
        if (!img) {
            fclose(spec);
            return seterr(err, errlen, "cannot write %s", name);
        }
        fprintf(img, "%d %d %e %e %s %e %e %e %e\n", W, H, jf * spectrum[f], m->freqs[f],
                m->GRMHD_FILE, m->M_UNIT, incl, m->R_LOW, m->R_HIGH);
        /* utilities.c:51-61: x outer, y inner, index x + y*W */
        for (int i = 0; i < W; i++)
            for (int j = 0; j < H; j++)
                fprintf(img, "%+.5e\n", jf * image[((size_t)i + (size_t)j * W) * nf + f]);
        fclose(img);
        fprintf(stderr, "Frequency %.5e Integrated flux density = %.5e\n", m->freqs[f],
                jf * spectrum[f]);
        fprintf(spec, "%+.15e\t%+.15e\n", m->freqs[f], jf * spectrum[f]);
    }
    fclose(spec);
    return 0;
}
