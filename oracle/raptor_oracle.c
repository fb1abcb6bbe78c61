/*
 * raptor_oracle.c -- FP64 CPU restatement of RAPTOR 1.0's per-pixel path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * raptor_amd/csrc; it is never linked into, loaded by or called from the
 * product.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load liboracle.so.
 *
 * It restates, function by function and in FP64 like the reference
 * (parameters.h:13, `typedef double real`), the path
 *   core.c:74-152 calculate_image
 *   integrator.c:44-70,109-146,149-289 (RK2, stepsize, f_geodesic, integrate_geodesic)
 *   metric.c:94-138,207-231,538-638,650-727 (MKS metric, inverse, connection, camera)
 *   GRmath.c:17-27,135-184 (lower_index, BL_to_KS_u)
 *   newtonraphson.c:11-97 (theta -> X2 inversion; solved deterministically here)
 *   radiative_transfer.c:17-119, j_nu.c:15,20-26,130-153,188-200
 *   j_nu.c:28-106,157-167 (kappa and angle-averaged thermal emission models)
 *   raptor_harm_model.c:190-320, grmhd.c:17-71 (fluid sampling, units)
 *   utilities.c:17-45 (cutoff_inner, R_GRAV, JANSKY_FACTOR)
 * with the reference's discretisation choices kept on purpose (truncating cell
 * lookup, end-point quadrature, tau gate, sigma cut, K2 ~ 2 Theta^2, gamma=4/3
 * in beta, nu_p floor, NaN poisoning when B_nu underflows).
 *
 * Differences from the reference, each deliberate:
 *  - no globals: every call takes an or_params; the routine is re-entrant and
 *    is run with OpenMP over pixels (the reference's pragma at core.c:105 is
 *    commented out and would race on Thetae_unit, raptor_harm_model.c:270).
 *  - the camera's theta -> X2 inversion is solved by Newton iteration from a
 *    fixed start to full convergence instead of from rcarry random restarts
 *    with a 1e-6 relative stop (newtonraphson.c:49-72); the root is unique.
 *  - all W*H pixels are traced (the reference skips pixels >= 1e6, core.c:81-103).
 *  - an optional tangent (forward-mode) pass carries d/d(ln M_unit) and
 *    d/d(R_high) of every accumulator; it does not change the image.
 *  - a 3D extension: N3 > 1 interpolates trilinearly, periodic in X3 = phi.
 *    With N3 == 1 the reference's 2D bilinear path is used unchanged.
 *
 * Primitive storage: prims[((v*n1 + i)*n2 + j)*n3 + k], v in the reference's
 * order KRHO, UU, U1, U2, U3, B1, B2, B3 (raptor_harm_model.h:67-74).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* constants.h:16-28, j_nu.c:15, parameters.h:144,170-185 */
#define OR_ELECTRON_CHARGE 4.8032068e-10
#define OR_ELECTRON_MASS 9.1093826e-28
#define OR_PROTON_MASS 1.67262171e-24
#define OR_BOLTZMANN 1.3806505e-16
#define OR_C 2.99792458e10
#define OR_PLANCK 6.6260693e-27
#define OR_GGRAV 6.6742e-8
#define OR_PI 3.14159265358979323846
#define OR_E2_C 7.695589048e-30
#define OR_E2_CH 1.161410e-03
#define OR_SOURCE_DIST 2.6228263e22
#define OR_RCAM 1.e4
#define OR_CUTOFF_OUTER (OR_RCAM * 1.01)
#define OR_HORIZON_MARG 1e-1
#define OR_MAX_ORDER 100

typedef struct {
    int n1, n2, n3;
    int width, height;
    int max_steps;            /* 1e5 in the reference (parameters.h:182) */
    double startx1, startx2, startx3;
    double dx1, dx2, dx3;
    double a, hslope, R0;
    double MBH, M_unit, R_high, R_low;
    double cam_size_x, cam_size_y;
    double inclination;       /* degrees */
    double stepsize;          /* STEPSIZE, model.in:12 */
    int integrator;           /* int_method (parameters.h:186-189): 0 RK2, 1 RK4, 2 Verlet */
    int emission;             /* 0 THSYNCH (the path), 1 THSYNCHAV, 2 kappa; see or_emission */
} or_params;

/* per-ray counters: RK2 steps, samples inside the grid gate, samples passing
 * the sigma cut (transfer actually evaluated); and per frequency the distance
 * of tau from ln 1000 at the sample where it crosses that gate
 * (radiative_transfer.c:90-95): min(ln 1000 - tau before, tau after - ln 1000),
 * or +inf if tau never crosses.  A transfer whose tau differs from this one
 * by more than that margin adds or drops one sample's e^-6.9-weighted term. */
typedef struct {
    int steps, n_grid, n_rt, escaped;
    int grid_step;   /* steps taken before the first one that lands inside the grid, or -1 */
    int opaque_step; /* steps taken when every frequency's tau had passed ln 1000, or -1:
                      * after it no sample changes I or its tangents */
    double gate[64];
} or_counters;

/* ---------------------------------------------------------------- units */
/* grmhd.c:17-33 set_units, utilities.c:17-21 set_constants */
typedef struct {
    double L_unit, RHO_unit, B_unit, Ne_unit, Rg, cutoff_inner;
    double stopx1, stopx2;
} or_derived;

static void or_derive(const or_params *P, or_derived *D)
{
    D->L_unit = OR_GGRAV * P->MBH / (OR_C * OR_C);
    D->RHO_unit = P->M_unit / pow(D->L_unit, 3);
    D->B_unit = OR_C * sqrt(4. * OR_PI * D->RHO_unit);
    D->Ne_unit = D->RHO_unit / (OR_PROTON_MASS + OR_ELECTRON_MASS);
    D->Rg = OR_GGRAV * P->MBH / OR_C / OR_C;
    D->cutoff_inner = (1. + sqrt(1. - P->a * P->a)) * (1. + OR_HORIZON_MARG);
    /* raptor_harm_model.c:126-128 */
    D->stopx1 = P->startx1 + P->n1 * P->dx1;
    D->stopx2 = P->startx2 + P->n2 * P->dx2;
}

double or_jansky_factor(const or_params *P)
{
    /* utilities.c:32-42 (LINEAR_IMPACT_CAM) */
    double Rg = OR_GGRAV * P->MBH / OR_C / OR_C;
    double dx = P->cam_size_x * Rg / OR_SOURCE_DIST;
    double dy = P->cam_size_y * Rg / OR_SOURCE_DIST;
    return 1.e23 * (dx / (double)P->width) * (dy / (double)P->height);
}

/* ------------------------------------------------------- MKS metric */
/* metric.c:94-138 */
void or_metric_dd(const or_params *P, const double X[4], double g[16])
{
    double a = P->a, h = P->hslope;
    memset(g, 0, 16 * sizeof(double));
    double r = exp(X[1]) + P->R0;
    double th = OR_PI * X[2] + 0.5 * (1. - h) * sin(2. * OR_PI * X[2]);
    double s = sin(th), s2 = s * s, c = cos(th);
    double rho2 = r * r + a * a * c * c;
    double rfac = r - P->R0;
    double hfac = OR_PI + (1. - h) * OR_PI * cos(2. * OR_PI * X[2]);
    double z = 2. * r / rho2;
    g[0 * 4 + 0] = -1. + z;
    g[0 * 4 + 1] = z * rfac;
    g[0 * 4 + 3] = -2. * a * r * s2 / rho2;
    g[1 * 4 + 0] = g[0 * 4 + 1];
    g[1 * 4 + 1] = (1. + z) * rfac * rfac;
    g[1 * 4 + 3] = -a * s2 * (1. + z) * rfac;
    g[2 * 4 + 2] = rho2 * hfac * hfac;
    g[3 * 4 + 0] = g[0 * 4 + 3];
    g[3 * 4 + 1] = g[1 * 4 + 3];
    g[3 * 4 + 3] = s2 * (rho2 + a * a * s2 * (1. + z));
}

/* metric.c:207-231 (closed form, valid for R0 = 0 as in the reference) */
void or_metric_uu(const or_params *P, const double X[4], double g[16])
{
    double a = P->a, h = P->hslope;
    memset(g, 0, 16 * sizeof(double));
    double r = exp(X[1]) + P->R0;
    double th = OR_PI * X[2] + 0.5 * (1. - h) * sin(2. * OR_PI * X[2]);
    double s = sin(th), s2 = s * s, c = cos(th);
    double irho2 = 1. / (r * r + a * a * c * c);
    double hfac = OR_PI + (1. - h) * OR_PI * cos(2. * OR_PI * X[2]);
    g[0 * 4 + 0] = -1. - 2. * r * irho2;
    g[0 * 4 + 1] = 2. * irho2;
    g[1 * 4 + 0] = g[0 * 4 + 1];
    g[1 * 4 + 1] = irho2 * (r * (r - 2.) + a * a) / (r * r);
    g[1 * 4 + 3] = a * irho2 / r;
    g[2 * 4 + 2] = irho2 / (hfac * hfac);
    g[3 * 4 + 1] = g[1 * 4 + 3];
    g[3 * 4 + 3] = irho2 / s2;
}

/* ------------------------------------------------ connection (MKS) */
/* metric.c:538-638.  G[(mu*4 + al)*4 + be] = Gamma^mu_{al be}. */
void or_connection(const or_params *P, const double X[4], double G[64])
{
    double a = P->a, h = P->hslope;
    double r = P->R0 + exp(X[1]);
    double r2 = r * r;
    double rp = exp(X[1]);
    double rp2 = rp * rp;
    double rpp = rp;
    double th = OR_PI * X[2] + 0.5 * (1. - h) * sin(2. * OR_PI * X[2]);
    double thp = OR_PI * (1. + (1. - h) * cos(2. * OR_PI * X[2]));
    double thp2 = thp * thp;
    double thpp = -2. * OR_PI * OR_PI * (1. - h) * sin(2. * OR_PI * X[2]);
    double c = cos(th), c2 = c * c, s = sin(th), s2 = s * s;
    double s2th = sin(2. * th);
    double cot = 1. / tan(th);
    double a2 = a * a;
    double Sig = r2 + a2 * c2;
    double Sig2 = Sig * Sig;
    double Del = r2 - 2. * r + a2;
    double A = Sig * Del + 2. * r * (r2 + a2);
    double iS = 1. / Sig, iS2 = iS * iS, iS3 = iS * iS * iS;
    double B = (2. * r2 - Sig) * iS3;
    double C = r * iS - a2 * B * s2;
#define GM(m, i, j) G[((m) * 4 + (i)) * 4 + (j)]
    GM(0, 0, 0) = 2. * r * B;
    GM(0, 0, 1) = (Sig + 2. * r) * B * rp;
    GM(0, 0, 2) = -a2 * r * s2th * iS2 * thp;
    GM(0, 0, 3) = -2. * a * r * B * s2;
    GM(0, 1, 1) = 2. * (Sig + r) * B * rp2;
    GM(0, 1, 2) = GM(0, 0, 2) * rp;
    GM(0, 1, 3) = -a * s2 * GM(0, 0, 1);
    GM(0, 2, 2) = -2. * r2 * iS * thp2;
    GM(0, 2, 3) = -a * s2 * GM(0, 0, 2);
    GM(0, 3, 3) = -2. * r * C * s2;

    GM(1, 0, 0) = Del * B / rp;
    GM(1, 0, 1) = (Del - Sig) * B;
    GM(1, 0, 2) = 0.;
    GM(1, 0, 3) = -a * Del * B * s2 / rp;
    GM(1, 1, 1) = rpp / rp - (2. * Sig - Del) * B * rp;
    GM(1, 1, 2) = -a2 * s * c * iS * thp;
    GM(1, 1, 3) = a * (r * iS + (Sig - Del) * B) * s2;
    GM(1, 2, 2) = -r * Del * iS / rp * thp2;
    GM(1, 2, 3) = 0.;
    GM(1, 3, 3) = -Del * C * s2 / rp;

    GM(2, 0, 0) = iS * GM(0, 0, 2) / thp2;
    GM(2, 0, 1) = GM(2, 0, 0) * rp;
    GM(2, 0, 2) = 0.;
    GM(2, 0, 3) = a * r * (r2 + a2) * s2th * iS3 / thp;
    GM(2, 1, 1) = GM(2, 0, 0) * rp2;
    GM(2, 1, 2) = r * iS * rp;
    GM(2, 1, 3) = a * s * c * (A + Sig * (Sig - Del)) * iS3 * rp / thp;
    GM(2, 2, 2) = thpp / thp + GM(1, 1, 2);
    GM(2, 2, 3) = 0.;
    GM(2, 3, 3) = -s * c * (Del * Sig2 + 2. * r * (r2 + a2) * (r2 + a2)) * iS3 / thp;

    GM(3, 0, 0) = a * B;
    GM(3, 0, 1) = GM(3, 0, 0) * rp;
    GM(3, 0, 2) = -2. * a * r * cot * iS2 * thp;
    GM(3, 0, 3) = -a2 * B * s2;
    GM(3, 1, 1) = GM(3, 0, 0) * rp2;
    GM(3, 1, 2) = -a * (Sig + 2. * r) * cot * iS2 * rp * thp;
    GM(3, 1, 3) = C * rp;
    GM(3, 2, 2) = -a * r * iS * thp2;
    GM(3, 2, 3) = (cot + a2 * r * s2th * iS2) * thp;
    GM(3, 3, 3) = -a * C * s2;
    for (int m = 0; m < 4; m++)
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < i; j++)
                GM(m, i, j) = GM(m, j, i);
#undef GM
}

/* ------------------------------------------------ geodesic stepping */
/* integrator.c:126-146 f_geodesic */
static void or_f_geodesic(const or_params *P, const double y[8], double f[8])
{
    double G[64];
    or_connection(P, y, G);
    const double *k = y + 4;
    double A[4] = {0., 0., 0., 0.};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            for (int l = 0; l < 4; l++)
                A[i] -= G[(i * 4 + j) * 4 + l] * k[j] * k[l];
    for (int i = 0; i < 4; i++) {
        f[i] = k[i];
        f[i + 4] = A[i];
    }
}

/* integrator.c:44-70 rk2_step (midpoint) */
void or_rk2_step(const or_params *P, double y[8], double dt)
{
    double ys[8], f[8], dx1[8];
    memcpy(ys, y, sizeof ys);
    or_f_geodesic(P, ys, f);
    for (int i = 0; i < 8; i++)
        ys[i] = y[i] + dt * f[i] * 0.5;
    or_f_geodesic(P, ys, f);
    for (int i = 0; i < 8; i++)
        dx1[i] = dt * f[i];
    for (int i = 0; i < 8; i++)
        y[i] = y[i] + dx1[i];
}

/* integrator.c:13-39 rk4_step: classical RK4 with weights (1/2, 1/2, 1) and
 * the update (dx0 + 2 dx1 + 2 dx2 + dx3) / 6 */
void or_rk4_step(const or_params *P, double y[8], double dt)
{
    const double w[4] = {0.5, 0.5, 1., 0.};
    double ys[8], f[8], dx[4][8];
    memcpy(ys, y, sizeof ys);
    for (int q = 0; q < 4; q++) {
        or_f_geodesic(P, ys, f);
        for (int i = 0; i < 8; i++) {
            dx[q][i] = dt * f[i];
            ys[i] = y[i] + dx[q][i] * w[q];
        }
    }
    for (int i = 0; i < 8; i++)
        y[i] = y[i] + 1. / 6. * (dx[0][i] + dx[1][i] * 2. + dx[2][i] * 2. + dx[3][i]);
}

/* integrator.c:71-105 verlet_step, velocity Verlet (Dolence et al. 2009
 * eq. 14a-d): X += dl k + dl^2 A / 2, k~ = k + dl A, then
 * k += dl (A(X, k) + A(X', k~)) / 2 */
void or_verlet_step(const or_params *P, double y[8], double dl)
{
    double ys[8], f[8], A0[4];
    memcpy(ys, y, sizeof ys);
    or_f_geodesic(P, ys, f);
    for (int i = 0; i < 4; i++) {
        ys[i] += dl * ys[i + 4] + 0.5 * dl * dl * f[i + 4];
        ys[i + 4] = ys[i + 4] + f[i + 4] * dl;
        A0[i] = f[i + 4];
    }
    or_f_geodesic(P, ys, f);
    for (int i = 0; i < 4; i++) {
        y[i] = ys[i];
        y[i + 4] += 0.5 * (A0[i] + f[i + 4]) * dl;
    }
}

/* integrator.c:258-263: the stepper chosen by int_method */
void or_step(const or_params *P, double y[8], double dl)
{
    if (P->integrator == 1)
        or_rk4_step(P, y, dl);
    else if (P->integrator == 2)
        or_verlet_step(P, y, dl);
    else
        or_rk2_step(P, y, dl);
}

/* integrator.c:109-122 stepsize */
double or_stepsize(const or_params *P, const double X[4], const double U[4])
{
    const double SMALL = 1.e-20;
    double dlx1 = P->stepsize / (fabs(U[1]) + SMALL * SMALL);
    double dlx2 = P->stepsize * fmin(X[2], 1. - X[2]) / (fabs(U[2]) + SMALL * SMALL);
    double dlx3 = P->stepsize / (fabs(U[3]) + SMALL * SMALL);
    double i1 = 1. / (fabs(dlx1) + SMALL * SMALL);
    double i2 = 1. / (fabs(dlx2) + SMALL * SMALL);
    double i3 = 1. / (fabs(dlx3) + SMALL * SMALL);
    return -fmax(1. / (i1 + i2 + i3), 1e-12);
}

/* -------------------------------------------------------------- camera */
/* newtonraphson.c:11-72: root of pi*X + (1-h)/2 sin(2 pi X) = theta.
 * f' = pi (1 + (1-h) cos 2 pi X) > 0 for h in (0,1], so the root is unique
 * and Newton from any start converges; iterate to a fixed point. */
double or_theta_to_x2(double theta, double h)
{
    double x = theta / OR_PI, prev = -1.;
    for (int it = 0; it < 200 && x != prev; it++) {
        prev = x;
        double f = OR_PI * x + 0.5 * (1. - h) * sin(2. * OR_PI * x) - theta;
        double fp = OR_PI + OR_PI * (1. - h) * cos(2. * OR_PI * x);
        x = x - f / fp;
        if (fabs(x - prev) <= 1e-15 * fabs(x))
            break;
    }
    return fmin(1., fmax(x, 0.));
}

/* metric.c:650-727 initialize_photon + GRmath.c:135-184 BL_to_KS_u +
 * newtonraphson.c:76-97 (U2 = U_theta / theta'(X2), linear) */
void or_init_photon(const or_params *P, double alpha, double beta, double y[8])
{
    double a = P->a;
    double mu0 = cos(P->inclination / 180. * OR_PI);
    double Xc[4] = {0., log(OR_RCAM), acos(mu0), 90. / 180. * OR_PI};
    double ll = -alpha * sqrt(1. - mu0 * mu0);
    double qq = beta * beta + mu0 * mu0 * (alpha * alpha - 1.);
    double kd[4], ku[4];
    double sinc2 = sin(Xc[2]) * sin(Xc[2]);
    double cosc2 = cos(Xc[2]) * cos(Xc[2]);
    double sgn = (beta < 0.) ? -1. : ((beta > 0.) ? 1. : 0.);
    kd[0] = -1.;
    kd[3] = ll;
    kd[2] = sgn * sqrt(fabs(qq - ll * ll * (cosc2 / sinc2) + cosc2));
    double r = OR_RCAM, rfactor = r;
    double theta = acos(mu0);
    double st = sin(theta), ct = cos(theta);
    double sigma = r * r + a * a * ct * ct;
    double delta = r * r + a * a - 2. * r;
    double A_ = (r * r + a * a) * (r * r + a * a) - delta * a * a * st * st;
    double g11 = sigma / delta * rfactor * rfactor;
    double g00 = -A_ / (sigma * delta);
    double g03 = -2. * a * r / (sigma * delta);
    double g33 = (delta - a * a * st * st) / (sigma * delta * st * st);
    double g22 = 1. / sigma;
    ku[0] = g00 * kd[0] + g03 * kd[3];
    ku[3] = g33 * kd[3] + g03 * kd[0];
    ku[2] = g22 * kd[2];
    ku[1] = sqrt((-ku[0] * kd[0] - ku[2] * kd[2] - ku[3] * kd[3]) / g11);

    /* BL -> KS (McKinney & Gammie 2004), log-radius form */
    double rc = exp(Xc[1]);
    double dc = rc * rc - 2. * rc + a * a;
    double t01 = 2. * rc / dc * rc;
    double t31 = a / dc * rc;
    double rp = 1. + sqrt(1. - a * a), rm = 1. - sqrt(1. - a * a);
    double lg = log((rc - rp) / (rc - rm));
    y[0] = Xc[0] + (log(dc) + 1. / sqrt(1. - a * a) * lg);
    y[1] = Xc[1];
    y[2] = Xc[2];
    y[3] = Xc[3] + (a / (2. * sqrt(1. - a * a)) * lg);
    y[4] = ku[0] + t01 * ku[1];
    y[5] = ku[1];
    y[6] = ku[2];
    y[7] = ku[3] + t31 * ku[1];
    /* BL theta -> MKS X2 */
    double x2 = or_theta_to_x2(y[2], P->hslope);
    y[2] = x2;
    y[6] = y[6] / (OR_PI * (1. + (1. - P->hslope) * cos(2. * OR_PI * x2)));
}

/* --------------------------------------------------------------- fluid */
typedef struct {
    double ne, thetae, B, mu_unused;
    double ucon[4], bcon[4];
    double dthetae_dRhigh;     /* tangent of Theta_e w.r.t. R_high */
    double beta2;
} or_fluid_t;

/* grmhd.c:46-71 Xtoij: C truncation toward zero, clamp with del = 0 / 1 */
static void or_xtoij(const or_params *P, const double X[4], int *i, int *j, double del[3])
{
    *i = (int)((X[1] - P->startx1) / P->dx1 - 0.5);
    *j = (int)((X[2] - P->startx2) / P->dx2 - 0.5);
    if (*i < 0) {
        *i = 0;
        del[1] = 0.;
    } else if (*i > P->n1 - 2) {
        *i = P->n1 - 2;
        del[1] = 1.;
    } else {
        del[1] = (X[1] - ((*i + 0.5) * P->dx1 + P->startx1)) / P->dx1;
    }
    if (*j < 0) {
        *j = 0;
        del[2] = 0.;
    } else if (*j > P->n2 - 2) {
        *j = P->n2 - 2;
        del[2] = 1.;
    } else {
        del[2] = (X[2] - ((*j + 0.5) * P->dx2 + P->startx2)) / P->dx2;
    }
}

/* The 3D extension's X3 index: periodic in phi with cell centres at
 * startx3 + (k + 1/2) dx3; linear weight del3 in [0,1). */
static void or_x3index(const or_params *P, double X3, int *k0, int *k1, double *del3)
{
    double period = P->n3 * P->dx3;
    double ph = fmod(X3 - P->startx3, period);
    if (ph < 0.)
        ph += period;
    double t = ph / P->dx3 - 0.5;
    double fk = floor(t);
    *del3 = t - fk;
    int k = (int)fk;
    *k0 = (k + P->n3) % P->n3;
    *k1 = (k + 1) % P->n3;
}

/* raptor_harm_model.c:190-320 get_fluid_params (+ grmhd.c:35-44 interp).
 * Returns 0 when b^2/rho > 1 (the sample is skipped), 1 otherwise. */
int or_fluid(const or_params *P, const or_derived *D, const double *prims,
             const double X[4], const double gcov[16], const double gcon[16], or_fluid_t *F)
{
    const double smalll = 1.e-20;
    int i, j;
    double del[3];
    or_xtoij(P, X, &i, &j, del);
    double c[4];
    c[0] = (1. - del[1]) * (1. - del[2]);
    c[1] = (1. - del[1]) * del[2];
    c[2] = del[1] * (1. - del[2]);
    c[3] = del[1] * del[2];
    double v[8];
    size_t n1 = P->n1, n2 = P->n2, n3 = P->n3;
    if (P->n3 == 1) {
        for (int q = 0; q < 8; q++) {
            const double *pv = prims + q * n1 * n2;
            v[q] = pv[i * n2 + j] * c[0] + pv[i * n2 + j + 1] * c[1] +
                   pv[(i + 1) * n2 + j] * c[2] + pv[(i + 1) * n2 + j + 1] * c[3];
        }
    } else {
        int k0, k1;
        double d3;
        or_x3index(P, X[3], &k0, &k1, &d3);
        for (int q = 0; q < 8; q++) {
            const double *pv = prims + q * n1 * n2 * n3;
#define PV(ii, jj, kk) pv[((size_t)(ii) * n2 + (jj)) * n3 + (kk)]
            double v0 = PV(i, j, k0) * c[0] + PV(i, j + 1, k0) * c[1] +
                        PV(i + 1, j, k0) * c[2] + PV(i + 1, j + 1, k0) * c[3];
            double v1 = PV(i, j, k1) * c[0] + PV(i, j + 1, k1) * c[1] +
                        PV(i + 1, j, k1) * c[2] + PV(i + 1, j + 1, k1) * c[3];
#undef PV
            v[q] = (1. - d3) * v0 + d3 * v1;
        }
    }
    double rho = v[0], uu = v[1];
    double Vcon[4] = {0., v[2], v[3], v[4]};
    double Bp[4] = {0., v[5], v[6], v[7]};
    F->ne = rho * D->Ne_unit;

    double VdotV = 0.;
    for (int a = 1; a < 4; a++)
        for (int b = 1; b < 4; b++)
            VdotV += gcov[a * 4 + b] * Vcon[a] * Vcon[b];
    double Vfac = sqrt(-1. / gcon[0] * (1. + fabs(VdotV)));
    F->ucon[0] = -Vfac * gcon[0];
    for (int a = 1; a < 4; a++)
        F->ucon[a] = Vcon[a] - Vfac * gcon[a];
    double ucov[4] = {0, 0, 0, 0};
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++)
            ucov[a] += gcov[a * 4 + b] * F->ucon[b];
    double UdotBp = 0.;
    for (int a = 1; a < 4; a++)
        UdotBp += ucov[a] * Bp[a];
    F->bcon[0] = UdotBp;
    for (int a = 1; a < 4; a++)
        F->bcon[a] = (Bp[a] + F->ucon[a] * UdotBp) / F->ucon[0];
    double bcov[4] = {0, 0, 0, 0};
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++)
            bcov[a] += gcov[a * 4 + b] * F->bcon[b];
    double Bsq = F->bcon[0] * bcov[0] + F->bcon[1] * bcov[1] + F->bcon[2] * bcov[2] +
                 F->bcon[3] * bcov[3];
    F->B = sqrt(Bsq) * D->B_unit + smalll;

    double beta = uu * (4. / 3. - 1.) / (0.5 * (Bsq + smalll) * 1.);
    double b2 = pow(uu * (4. / 3. - 1.) / (0.5 * (Bsq + smalll) * 1.), 2);
    (void)beta;
    F->beta2 = b2;
    double trat = P->R_high * b2 / (1. + b2) + P->R_low / (1. + b2);
    double gam2 = 0.5 * ((1. + 2. / 3. * (trat + 1.) / (trat + 2.)) + 4. / 3.);
    double mr = OR_PROTON_MASS / OR_ELECTRON_MASS;
    double the_unit = (gam2 - 1.) * mr / (1. + trat);
    F->thetae = (uu / rho) * the_unit;
    /* tangent w.r.t. R_high: dtrat/dRh = b2/(1+b2) */
    double dgam2 = (1. / 3.) / ((trat + 2.) * (trat + 2.));
    double dunit = mr * (dgam2 * (1. + trat) - (gam2 - 1.)) / ((1. + trat) * (1. + trat));
    F->dthetae_dRhigh = (uu / rho) * dunit * (b2 / (1. + b2));
    if (F->thetae < 0) {
        F->thetae = smalll;
        F->dthetae_dRhigh = 0.;
    }
    if (Bsq / rho > 1.0) {
        return 0;
    }
    return 1;
}

/* ------------------------------------------------------------ emission */
/* j_nu.c:130-153 emission_coeff_THSYNCH (mu = cos pitch angle) */
double or_j_thsynch(double B, double mu, double Te, double nu_p, double ne)
{
    double sth = sqrt(1 - mu * mu);
    double nu_c = OR_ELECTRON_CHARGE * B / (2. * OR_PI * OR_ELECTRON_MASS * OR_C);
    double nu_s = (2. / 9.) * nu_c * Te * Te * sth;
    double X = nu_p / (nu_s);
    double f = pow(pow(X, 0.5) + pow(2., 11. / 12.) * pow(X, 1. / 6.), 2.);
    if (nu_p == 1 || fabs(sth) < 1.e-10)
        return 0;
    double K2;
    if (Te > 1e-4)
        K2 = 2 * Te * Te;
    else
        return 0;
    double j = ne * sqrt(2.) * OR_PI * nu_s / (3. * K2) * f * exp(-pow(X, 1. / 3.));
    return j * OR_E2_C;
}

/* j_nu.c:188-191 planck_function (no factor h; compensated by e2_ch) */
double or_planck(double nu, double Te)
{
    double T = Te * OR_ELECTRON_MASS * OR_C * OR_C / OR_BOLTZMANN;
    return 2. * nu * nu * nu / (OR_C * OR_C) * 1. / (exp(OR_PLANCK * nu / (OR_BOLTZMANN * T)) - 1.);
}

/* j_nu.c:194-200 absorption_coeff_TH (Kirchhoff) */
double or_a_th(double j, double nu, double Te)
{
    double Bn = or_planck(nu, Te);
    return j * OR_E2_CH / Bn / OR_E2_C;
}

/* ---------------------------------------------- other emission models */
/* Gauss hypergeometric series sum_n (a)_n (b)_n / ((c)_n n!) x^n, 0 <= x <= 1/2. */
static double or_2f1_series(double a, double b, double c, double x)
{
    double term = 1., sum = 1.;
    for (int n = 0; n < 2000; n++) {
        term *= (a + n) * (b + n) / ((c + n) * (n + 1.)) * x;
        sum += term;
        if (fabs(term) <= 1e-17 * fabs(sum))
            break;
    }
    return sum;
}

/* 2F1(a, b; c; z) for z <= 0 (the kappa absorptivity's z = -kappa Theta_e),
 * standing in for gsl_sf_hyperg_2F1 (GSL, version unpinned, absent here) at
 * j_nu.c:81-91.  -1 <= z <= 0: Pfaff's transformation (Abramowitz & Stegun
 * 15.3.4) 2F1(a, b; c; z) = (1-z)^-a 2F1(a, c-b; c; z/(z-1)), argument in
 * [0, 1/2]; z < -1: the continuation in 1/(1-z) (A&S 15.3.8) exactly as
 * j_nu.c:81-88 writes it, the inner series' argument in (0, 1/2).  b - a must
 * not be an integer (it is 4/3 for the kappa model). */
double or_hyp2f1(double a, double b, double c, double z)
{
    if (z >= -1.)
        return pow(1. - z, -a) * or_2f1_series(a, c - b, c, z / (z - 1.));
    const double x = 1. / (1. - z);
    return pow(1. - z, -a) * tgamma(c) * tgamma(b - a) / (tgamma(b) * tgamma(c - a)) *
               or_2f1_series(a, c - b, a - b + 1., x) +
           pow(1. - z, -b) * tgamma(c) * tgamma(a - b) / (tgamma(a) * tgamma(c - b)) *
               or_2f1_series(b, c - a, b - a + 1., x);
}

/* j_nu.c:28-61 emission_coeff_kappa_FIT (Pandya et al. 2016), kappa = 3.5,
 * width w = Theta_e; mu = cos pitch angle (the reference's `theta`).  cgs. */
double or_j_kappa(double nu, double Ne, double Te, double B, double mu)
{
    const double kappa = 3.5, w = Te;
    double nuc = OR_ELECTRON_CHARGE * B / (2. * OR_PI * OR_ELECTRON_MASS * OR_C);
    double sth = sqrt(1 - mu * mu);
    double factor = (Ne * pow(OR_ELECTRON_CHARGE, 2.) * nuc * sth) / OR_C;
    double nus = nuc * sth * pow(w * kappa, 2);
    if (nu == 1.0 || fabs(mu) == 1)
        return 0;
    double X = nu / nus;
    double J_low = pow(X, 1. / 3.) * 4 * OR_PI * tgamma(kappa - 4. / 3.) /
                   (pow(3, 7. / 3.) * tgamma(kappa - 2.));
    double J_high = pow(X, -(kappa - 2) / 2.) * pow(3, (kappa - 1) / 2.) * (kappa - 2.) *
                    (kappa - 1.) / 4. * tgamma(kappa / 4. - 1. / 3.) *
                    tgamma(kappa / 4. + 4. / 3.);
    double x = 3 * pow(kappa, -3. / 2.);
    double J_s = pow((pow(J_low, -x) + pow(J_high, -x)), -1. / x);
    return J_s * factor;
}

/* j_nu.c:63-106 absorption_coeff_kappa_FIT (Pandya et al. 2016).  cgs. */
double or_a_kappa(double nu, double Ne, double Te, double B, double mu)
{
    const double kappa = 3.5, w = Te;
    double nuc = OR_ELECTRON_CHARGE * B / (2. * OR_PI * OR_ELECTRON_MASS * OR_C);
    double sth = sqrt(1 - mu * mu);
    double factor = Ne * OR_ELECTRON_CHARGE / (B * sth);
    if (nu == 1.0 || fabs(mu) == 1.)
        return 0;
    double nus = nuc * sth * pow(w * kappa, 2);
    double X = nu / nus;
    double a = kappa - 1. / 3., b = kappa + 1., c = kappa + 2. / 3., z = -kappa * w;
    if (fabs(z) == 1.)
        return 0;
    double hyp = or_hyp2f1(a, b, c, z);
    double A_low = pow(X, -5. / 3.) * pow(3, 1. / 6.) * (10. / 41.) * pow(2 * OR_PI, 2) /
                   pow(w * kappa, 16. / 3. - kappa) * (kappa - 2.) * (kappa - 1.) * kappa /
                   (3. * kappa - 1.) * tgamma(5. / 3.) * hyp;
    double A_high = pow(X, -(3. + kappa) / 2.) * (pow(OR_PI, 5. / 2.) / 3.) *
                    ((kappa - 2.) * (kappa - 1.) * kappa / pow(w * kappa, 5.)) *
                    (2 * tgamma(2. + kappa / 2.) / (2. + kappa) - 1.) *
                    (pow(3. / kappa, 19. / 4.) + 3. / 5.);
    double x = pow(-7. / 4. + 8. * kappa / 5., -43. / 50.);
    double A_s = pow((pow(A_low, -x) + pow(A_high, -x)), -1. / x);
    return factor * A_s;
}

/* j_nu.c:157-167 emission_coeff_THSYNCHAV (Dexter 2009, angle-averaged), as
 * written: without the e^2/c that emission_coeff_THSYNCH multiplies by. */
double or_j_thsynchav(double B, double Te, double nu_p, double ne)
{
    double nu_c = 3. * OR_ELECTRON_CHARGE * B * Te * Te / (4. * OR_PI * OR_ELECTRON_MASS * OR_C);
    double x_M = nu_p / nu_c;
    double I = 4.0505 / pow(x_M, 1. / 6.) * (1. + 0.4 / pow(x_M, 1. / 4.) + 0.5316 / sqrt(x_M)) *
               exp(-1.8899 * pow(x_M, 1. / 3.));
    return nu_p * ne / (2. * sqrt(3.)) * 1. / (Te * Te) * I;
}

/* The emission model of a sample: emission_coeff (j_nu.c:20-26) calls the
 * thermal synchrotron fit and Kirchhoff's absorptivity; the reference defines
 * the other two models but calls neither.  Here they can be chosen:
 *   1  THSYNCHAV: j = e^2/c * emission_coeff_THSYNCHAV, in cgs like the
 *      thermal fit, so that absorption_coeff_TH (Kirchhoff) applies as is;
 *   2  kappa: emissivity and absorptivity fits of j_nu.c:28-106. */
void or_emission(int model, double B, double mu, double Te, double nu_p, double ne, double *j,
                 double *an)
{
    if (model == 2) {
        *j = or_j_kappa(nu_p, ne, Te, B, mu);
        *an = or_a_kappa(nu_p, ne, Te, B, mu);
    } else if (model == 1) {
        *j = OR_E2_C * or_j_thsynchav(B, Te, nu_p, ne);
        *an = or_a_th(*j, nu_p, Te);
    } else {
        *j = or_j_thsynch(B, mu, Te, nu_p, ne);
        *an = or_a_th(*j, nu_p, Te);
    }
}

/* d ln j / d ln X of the THSYNCH fit at fixed n_e, nu_s:
 * ln f = 2 ln(X^1/2 + 2^(11/12) X^1/6), minus X^1/3. */
static double or_dlnj_dlnX(double X)
{
    double c = pow(2., 11. / 12.);
    double s = pow(X, 0.5), t = pow(X, 1. / 6.);
    return 2. * (0.5 * s + c / 6. * t) / (s + c * t) - pow(X, 1. / 3.) / 3.;
}

/* radiative_transfer.c:88-96: one sample of the end-point quadrature.
 * tau accumulates dtau = nu_p*alpha*dl; dI = j/nu_p^2 exp(-tau) dl is added
 * only while 0 <= tau < ln 1000 (a NaN tau closes the gate for good). */
void or_transfer_step(double *tau, double *I, double j, double an, double nu_p, double dlf,
                      double *dtau_out, double *dI_out)
{
    double dtau = (nu_p * an * dlf);
    *tau += dtau;
    double dI = (j / (nu_p * nu_p)) * exp(-*tau) * dlf;
    if (*tau < log(1000.) && !(*tau < 0.0))
        *I += dI;
    *dtau_out = dtau;
    *dI_out = dI;
}

/* ------------------------------------------- plasma-frame quantities */
/* GRmath.c:17-27 lower_index: k_a = g_ab k^b */
static void or_lower(const double gcov[16], const double k[4], double kd[4])
{
    for (int a = 0; a < 4; a++) {
        kd[a] = 0.;
        for (int b = 0; b < 4; b++)
            kd[a] += gcov[a * 4 + b] * k[b];
    }
}

/* radiative_transfer.c:32-47 pitch_angle (the cosine), with k lowered as at
 * radiative_transfer.c:66: (k_a b^a) / (|k_a u^a| |b|), |b| = B / B_unit,
 * clamped to +-1 by mu / |mu|; 0 for B == 0. */
static double or_pitch_cos(const double gcov[16], const double k[4], const or_fluid_t *F,
                           double B_unit)
{
    if (F->B == 0.)
        return 0.;
    double kd[4];
    or_lower(gcov, k, kd);
    double kk = fabs(kd[0] * F->ucon[0] + kd[1] * F->ucon[1] + kd[2] * F->ucon[2] +
                     kd[3] * F->ucon[3]);
    double mu = (kd[0] * F->bcon[0] + kd[1] * F->bcon[1] + kd[2] * F->bcon[2] +
                 kd[3] * F->bcon[3]) / (kk * F->B / B_unit);
    if (fabs(mu) > 1.)
        mu /= fabs(mu);
    return mu;
}

/* radiative_transfer.c:17-29 freq_in_plasma_frame with the wave vector scaled
 * to frequency nu and lowered (radiative_transfer.c:74-79):
 * nu_p = -(u^a k_a) m_e c^2 / h with k = k_geo h nu / (m_e c^2); nu_p <= 0
 * becomes 1 Hz. */
static double or_nu_plasma(const double gcov[16], const double k[4], const double ucon[4],
                           double nu)
{
    double kuf[4], kdf[4];
    for (int a = 0; a < 4; a++)
        kuf[a] = k[a] * OR_PLANCK * nu / (OR_ELECTRON_MASS * OR_C * OR_C);
    or_lower(gcov, kuf, kdf);
    double nu_p = 0.;
    for (int a = 0; a < 4; a++)
        nu_p += ucon[a] * kdf[a];
    nu_p *= -(OR_ELECTRON_MASS * OR_C * OR_C) / OR_PLANCK;
    if (nu_p <= 0.0)
        nu_p = 1.0;
    return nu_p;
}

/* ------------------------------------------------------------ one ray */
/* integrator.c:149-289 integrate_geodesic + radiative_transfer.c:49-119.
 * I[f] returns the invariant intensity I_nu/nu^3 before the nu^3 scaling of
 * core.c:136; dI (if not NULL) returns [f][2] tangents (ln M_unit, R_high). */
void or_trace_ray(const or_params *P, const double *prims, int x, int yy, int nf,
                  const double *freqs, double *I, double *dI, or_counters *cnt)
{
    or_derived D;
    or_derive(P, &D);
    double tau[64], dtau_t[64][2];
    for (int f = 0; f < nf; f++) {
        tau[f] = 0.;
        I[f] = 0.;
        dtau_t[f][0] = dtau_t[f][1] = 0.;
        if (dI) dI[2 * f] = dI[2 * f + 1] = 0.;
    }
    double stepx = P->cam_size_x / (double)P->width;
    double stepy = P->cam_size_y / (double)P->height;
    double alpha = -P->cam_size_x * 0.5 + (x + 0.5) * stepx;
    double beta = -P->cam_size_y * 0.5 + (yy + 0.5) * stepy;
    double y[8];
    or_init_photon(P, alpha, beta, y);
    double r_current = exp(y[1]);
    int steps = 0, TERMINATE = 0, theta_turns = 0;
    double thetadot_prev = 0.;
    int n_grid = 0, n_rt = 0, escaped = 0, grid_step = -1, opaque_step = -1;
    const double lnth = log(1000.);
    double gate[64];
    for (int f = 0; f < nf; f++)
        gate[f] = INFINITY;

    while (r_current < OR_CUTOFF_OUTER && r_current > D.cutoff_inner &&
           steps < P->max_steps && !TERMINATE) {
        if (thetadot_prev * y[6] < 0. && steps > 2)
            theta_turns += 1;
        thetadot_prev = y[6];
        if ((beta < 0. && theta_turns > OR_MAX_ORDER) ||
            (beta > 0. && theta_turns > (OR_MAX_ORDER + 1)))
            TERMINATE = 1;
        double dl = or_stepsize(P, y, y + 4);
        double dlabs = fabs(dl);
        or_step(P, y, dl);
        if (y[1] > D.stopx1 && y[5] < 0) {
            escaped = 1;
            break;
        }
        r_current = exp(y[1]);
        if (y[1] < D.stopx1 && y[1] > P->startx1 && y[2] > P->startx2 && y[2] < D.stopx2) {
            if (grid_step < 0)
                grid_step = steps;
            n_grid++;
            double gcov[16], gcon[16];
            or_metric_dd(P, y, gcov);
            or_metric_uu(P, y, gcon);
            or_fluid_t F;
            if (or_fluid(P, &D, prims, y, gcov, gcon, &F)) {
                n_rt++;
                const double *k = y + 4;
                const double mu = or_pitch_cos(gcov, k, &F, D.B_unit);
                for (int f = 0; f < nf; f++) {
                    const double nu_p = or_nu_plasma(gcov, k, F.ucon, freqs[f]);
                    double j, an;
                    or_emission(P->emission, F.B, mu, F.thetae, nu_p, F.ne, &j, &an);
                    double dlf = dlabs * D.Rg / freqs[f];
                    double dtau, dIk;
                    const double tau0 = tau[f];
                    or_transfer_step(&tau[f], &I[f], j, an, nu_p, dlf, &dtau, &dIk);
                    if (tau0 < lnth && tau[f] >= lnth)
                        gate[f] = fmin(lnth - tau0, tau[f] - lnth);
                    /* tangents: d ln j and d ln alpha per parameter */
                    double dlnj[2] = {0., 0.}, dlna[2] = {0., 0.};
                    if (dI && j != 0.) {
                        double sth = sqrt(1 - mu * mu);
                        double nu_c = OR_ELECTRON_CHARGE * F.B / (2. * OR_PI * OR_ELECTRON_MASS * OR_C);
                        double nu_s = (2. / 9.) * nu_c * F.thetae * F.thetae * sth;
                        double Dx = or_dlnj_dlnX(nu_p / nu_s);
                        double dlnB = 0.5 * (F.B - 1e-20) / F.B;
                        double dlnT = F.dthetae_dRhigh / F.thetae;
                        double xx = OR_PLANCK * nu_p / (F.thetae * OR_ELECTRON_MASS * OR_C * OR_C);
                        double dlnBnu = xx / (-expm1(-xx));
                        dlnj[0] = 1. + dlnB * (1. - Dx);
                        dlnj[1] = -2. * Dx * dlnT;
                        dlna[0] = dlnj[0];
                        dlna[1] = dlnj[1] - dlnBnu * dlnT;
                    }
                    if (dI) {
                        for (int q = 0; q < 2; q++)
                            dtau_t[f][q] += (dtau != 0.) ? dtau * dlna[q] : 0.;
                    }
                    if (dI && tau[f] < lnth && !(tau[f] < 0.0)) {
                        for (int q = 0; q < 2; q++)
                            dI[2 * f + q] += (dIk != 0.) ? dIk * (dlnj[q] - dtau_t[f][q]) : 0.;
                    }
                }
                if (opaque_step < 0) {
                    int f = 0;
                    while (f < nf && tau[f] >= lnth)
                        f++;
                    if (f == nf)
                        opaque_step = steps;
                }
            }
        }
        steps = steps + 1;
    }
    if (cnt) {
        cnt->steps = steps;
        cnt->n_grid = n_grid;
        cnt->n_rt = n_rt;
        cnt->escaped = escaped;
        cnt->grid_step = grid_step;
        cnt->opaque_step = opaque_step;
        for (int f = 0; f < nf; f++)
            cnt->gate[f] = gate[f];
    }
}

/* core.c:74-152 calculate_image over a pixel list.
 * image[p*nf + f] = I_nu (after the nu^3 scaling of core.c:136) for pixel
 * pix[p] (index y*W + x); dimage (optional) [p][f][2]; counters optional [p][6];
 * gate (optional) [p][f]: the tau-gate margin of or_counters.
 * nthreads <= 0 uses the OpenMP default. */
int or_render(const or_params *P, const double *prims, int nf, const double *freqs,
              const int *pix, int npix, double *image, double *dimage, int *counters,
              double *gate, int nthreads)
{
    if (nf < 1 || nf > 64 || npix < 0)
        return -1;
    if (dimage && P->emission != 0)   /* the tangents are the thermal fit's */
        return -1;
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 4)
    for (int p = 0; p < npix; p++) {
        int i = pix ? pix[p] : p;
        int yy = i / P->width, x = i % P->width;
        double I[64], dI[128];
        or_counters c;
        or_trace_ray(P, prims, x, yy, nf, freqs, I, dimage ? dI : NULL, &c);
        for (int f = 0; f < nf; f++) {
            double nu3 = pow(freqs[f], 3.);
            image[(size_t)p * nf + f] = I[f] * nu3;
            if (dimage) {
                dimage[((size_t)p * nf + f) * 2 + 0] = dI[2 * f] * nu3;
                dimage[((size_t)p * nf + f) * 2 + 1] = dI[2 * f + 1] * nu3;
            }
        }
        if (counters) {
            counters[p * 6 + 0] = c.steps;
            counters[p * 6 + 1] = c.n_grid;
            counters[p * 6 + 2] = c.n_rt;
            counters[p * 6 + 3] = c.escaped;
            counters[p * 6 + 4] = c.grid_step;
            counters[p * 6 + 5] = c.opaque_step;
        }
        if (gate)
            for (int f = 0; f < nf; f++)
                gate[(size_t)p * nf + f] = c.gate[f];
    }
    return 0;
}

/* units exported for tests (grmhd.c:17-33, utilities.c:17-21) */
void or_units(const or_params *P, double out[6])
{
    or_derived D;
    or_derive(P, &D);
    out[0] = D.L_unit;
    out[1] = D.RHO_unit;
    out[2] = D.B_unit;
    out[3] = D.Ne_unit;
    out[4] = D.Rg;
    out[5] = D.cutoff_inner;
}

/* fluid sample exported for tests: out = {ne, thetae, B, ucon[4], bcon[4],
 * dthetae_dRhigh, ok} */
int or_fluid_at(const or_params *P, const double *prims, const double X[4], double out[13])
{
    or_derived D;
    or_derive(P, &D);
    double gcov[16], gcon[16];
    or_metric_dd(P, X, gcov);
    or_metric_uu(P, X, gcon);
    or_fluid_t F;
    int ok = or_fluid(P, &D, prims, X, gcov, gcon, &F);
    out[0] = F.ne;
    out[1] = F.thetae;
    out[2] = F.B;
    for (int a = 0; a < 4; a++) {
        out[3 + a] = F.ucon[a];
        out[7 + a] = F.bcon[a];
    }
    out[11] = F.dthetae_dRhigh;
    out[12] = ok;
    return ok;
}

/* one transfer sample for tests, composed as or_trace_ray composes it
 * (radiative_transfer.c:61-84): at y = (X, k) the fluid state, the pitch
 * angle's cosine and, per frequency, the plasma-frame frequency and the
 * emission coefficients.  out = {ok, ne, thetae, B, mu, then per frequency
 * nu_p, j_nu, alpha_nu}; ok = 0 where the sigma cut skips the sample (the
 * rest is then as computed, for inspection). */
int or_sample_at(const or_params *P, const double *prims, const double y[8], int nf,
                 const double *freqs, double *out)
{
    or_derived D;
    or_derive(P, &D);
    double gcov[16], gcon[16];
    or_metric_dd(P, y, gcov);
    or_metric_uu(P, y, gcon);
    or_fluid_t F;
    int ok = or_fluid(P, &D, prims, y, gcov, gcon, &F);
    const double mu = or_pitch_cos(gcov, y + 4, &F, D.B_unit);
    out[0] = ok;
    out[1] = F.ne;
    out[2] = F.thetae;
    out[3] = F.B;
    out[4] = mu;
    for (int f = 0; f < nf; f++) {
        const double nu_p = or_nu_plasma(gcov, y + 4, F.ucon, freqs[f]);
        double j, an;
        or_emission(P->emission, F.B, mu, F.thetae, nu_p, F.ne, &j, &an);
        out[5 + 3 * f] = nu_p;
        out[6 + 3 * f] = j;
        out[7 + 3 * f] = an;
    }
    return ok;
}

/* one ray's path for tests: writes up to maxpts states y[8] after each step
 * (including the initial state) and returns the number written */
int or_ray_path(const or_params *P, double alpha, double beta, double *out, int maxpts)
{
    or_derived D;
    or_derive(P, &D);
    double y[8];
    or_init_photon(P, alpha, beta, y);
    int n = 0;
    if (n < maxpts)
        memcpy(out + 8 * n++, y, sizeof y);
    double r_current = exp(y[1]);
    int steps = 0;
    while (r_current < OR_CUTOFF_OUTER && r_current > D.cutoff_inner && steps < P->max_steps) {
        double dl = or_stepsize(P, y, y + 4);
        or_step(P, y, dl);
        if (n < maxpts)
            memcpy(out + 8 * n++, y, sizeof y);
        if (y[1] > D.stopx1 && y[5] < 0)
            break;
        r_current = exp(y[1]);
        steps++;
    }
    return n;
}
