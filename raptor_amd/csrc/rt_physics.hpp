// rt_physics.hpp -- per-ray physics of the MI355X tracer (device code).
//
// One ray lives in one lane's registers: its state y = (X^mu, k^mu) in FP64,
// the trigonometry of its current point (shared by the metric of the
// transfer sample and the first RK2 stage of the next step), and per
// frequency the optical depth and intensity accumulators.
//
// The functions restate the reference's FP64 path (parameters.h:13):
//   camera            metric.c:650-727, GRmath.c:135-184, newtonraphson.c:76-97
//   step size         integrator.c:109-122
//   RK2 midpoint      integrator.c:44-70 (RK4 and velocity Verlet: 13-105)
//   geodesic accel.   integrator.c:126-146 with metric.c:538-638 (MKS Gamma)
//   metric            metric.c:94-138, 207-231
//   fluid sample      raptor_harm_model.c:190-320, grmhd.c:35-71
//   transfer          radiative_transfer.c:17-119
//   emissivity        j_nu.c:15, 130-153, 188-200 (thermal synchrotron);
//                     28-106 (kappa = 3.5), 157-167 (angle-averaged thermal)
// The connection's contraction with k is formed from the Kerr-Schild
// structure of the metric instead of filling a 4x4x4 array, and the
// trigonometry of a point (exp X1, sin/cos theta, sin/cos 2 pi X2) is
// computed once and carried from step to step.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <cmath>
#include <vector>

namespace rt {

// constants.h:16-28, j_nu.c:15, parameters.h:144,170-185
constexpr double kPi = 3.14159265358979323846;
constexpr double kElectronCharge = 4.8032068e-10;
constexpr double kElectronMass = 9.1093826e-28;
constexpr double kProtonMass = 1.67262171e-24;
constexpr double kBoltzmann = 1.3806505e-16;
constexpr double kC = 2.99792458e10;
constexpr double kPlanck = 6.6260693e-27;
constexpr double kGGrav = 6.6742e-8;
constexpr double kE2C = 7.695589048e-30;
constexpr double kE2CH = 1.161410e-03;
constexpr double kSourceDist = 2.6228263e22;
constexpr double kRcam = 1.e4;
constexpr double kCutoffOuter = kRcam * 1.01;
constexpr double kHorizonMarg = 1e-1;
constexpr int kMaxOrder = 100;
constexpr double kSmall = 1.e-20;
constexpr int kMaxFreq = 16;
constexpr int kMaxCam = 64;

// Per-inclination camera constants (the parts of initialize_photon that do
// not depend on the pixel), computed on the host.
struct Camera {
    double sin_i;            // sqrt(1 - mu0^2)
    double mu0sq;            // mu0^2
    double cot2;             // cos^2/sin^2 of the camera theta
    double cosc2;            // cos^2 of the camera theta
    double g00, g03, g33, g22, g11;   // BL (log-radius) metric at the camera
    double t01, t31;         // BL -> KS transform of k^t, k^phi
    double X1, X2, X3;       // camera position (MKS); X0 is never used
    double thp_inv;          // 1 / theta'(X2) at the camera
};

// Everything a trace launch reads; passed by value as the kernel argument.
// Emission models (rt_params.emission).  emission_coeff (j_nu.c:20-26) calls
// the thermal synchrotron fit; the reference also defines an angle-averaged
// thermal fit and the kappa-distribution fits, which nothing on its path
// calls, and which can be chosen here.
constexpr int kThSynch = 0, kThSynchAv = 1, kKappa = 2;

// Constants of the kappa = 3.5 fits (j_nu.c:28-106), computed on the host in
// FP64 (kappa_const) and used in FP32 in log2 form, so that the power laws
// stay in range down to the Theta_e floor of 1e-20.
struct KappaConst {
    float lJlow, lJhigh;   // log2 of J_low / X^(1/3) and J_high / X^(-(kappa-2)/2)
    float xj;              // J_s = (J_low^-xj + J_high^-xj)^(-1/xj)
    float lAlow, lAhigh;   // log2 of the constant factors of A_low and A_high
    float xa;              // A_s exponent
    float ha, hb, hc;      // hypergeometric parameters kappa - 1/3, kappa + 1, kappa + 2/3
    float g1, g2;          // Gamma-function ratios of the 1/(1-z) continuation
    // log2 of the absorptivity's hypergeometric factor tabulated over
    // u = log2(-z) (value and d/du at u0 + i / idu, i < n), built on the host
    // in FP64 (kappa_table); NULL evaluates the series instead
    const float2 *tab;
    float tu0, tidu;
    int tn;
};

// Largest grid the samplers address: cell offsets into the staged
// primitives (two float4, 32 bytes, per cell) are unsigned 32-bit byte
// offsets, 2^27 cells (4 GiB of FP32 primitives, e.g. 512^3).
constexpr long long kMaxCells = 1LL << 27;

struct Params {
    // grid (raptor_harm_model.c:120-135)
    int n1, n2, n3;
    unsigned s1, s2;             // float4 elements per step in i (2 n2 n3) and in j (2 n3)
    double startx1, startx2, startx3;
    double dx1, dx2, dx3;
    double idx1, idx2, idx3;     // 1 / dx
    double period3, iperiod3;    // n3 dx3 and its reciprocal (the X3 period)
    double stopx1, stopx2;
    double a, hslope;
    // camera and integrator
    int width, height, npix;
    double cam_x, cam_y;
    double stepsize;
    int max_steps;
    double cutoff_inner;
    // units (grmhd.c:17-33) and the electron model
    double Ne_unit, B_unit, Rg;
    double R_high, R_low;
    int nf;
    double freq[kMaxFreq];
    // FP32 copies for the transfer
    float fnu[kMaxFreq];
    float flnu6[kMaxFreq];       // (log2(freq) - 37) / 6
    float fkap[kMaxFreq];        // (e^2 / (c h)) c^2 / (2 freq^3)
    float fNe_unit, fB_unit, fiB_unit, fR_high, fR_low, fa, fpi_h;   // fpi_h = (1 - h) pi
    int emission;                // kThSynch, kThSynchAv or kKappa
    KappaConst kc;
    int ncam;
    Camera cam[kMaxCam];
};

// Trigonometry of one point.
struct Point {
    double r;      // exp(X1) (R0 = 0)
    double s, c;   // sin, cos of theta(X2)
    double s2, c2; // sin, cos of 2 pi X2
};

// Single-instruction FP32 reciprocal, square root, exp2 and log2 (about 1
// ulp) for the transfer pass, whose error budget is 1e-3 per pixel.  The
// __host__ overloads serve CPU builds of these functions for testing.
__device__ inline float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ inline float flog2(float x) { return __builtin_amdgcn_logf(x); }
__host__ inline float frcp(float x) { return 1.f / x; }
__host__ inline float fsqrt(float x) { return sqrtf(x); }
__host__ inline float fexp2(float x) { return exp2f(x); }
__host__ inline float flog2(float x) { return log2f(x); }
__host__ __device__ inline float fexp(float x) { return fexp2(1.4426950408889634f * x); }

// Whether any lane of the wave holds c: a wave-uniform branch condition, so
// that the branch is skipped rather than evaluated and selected away.
__device__ inline bool wave_any(bool c) { return __ballot(c) != 0ull; }
__host__ inline bool wave_any(bool c) { return c; }

// e^x - 1 for the Planck factor: the Taylor series to x^6/6! for |x| <= 1/4
// (truncation below x^6/7! < 5e-8 relative) and exp(x) - 1 beyond (the
// single-instruction exp2's ~1 ulp, at most 5.4e-7 relative at |x| = 1/4).
// The exp side is a branch, skipped by a wave none of whose lanes needs it:
// x = h nu / (k T_e) stays below 1e-4 for radio to sub-mm frequencies
// (up to 1e12 Hz) at every temperature the emissivity gates let through
// (Theta_e > 1e-4).
__host__ __device__ inline float fexpm1(float x)
{
    float p = 1.f / 720.f;
    p = p * x + 1.f / 120.f;
    p = p * x + 1.f / 24.f;
    p = p * x + 1.f / 6.f;
    p = p * x + 0.5f;
    p = p * x + 1.f;
    const float t = x * p;
    const bool far = !(fabsf(x) <= 0.25f);
    if (wave_any(far)) {
        asm volatile("");   // keeps the exp inside the branch (not speculated)
        const float e = fexp(x) - 1.f;
        return far ? e : t;
    }
    return t;
}

// FP64 reciprocal for the geodesic step: the hardware estimate refined by two
// Newton steps r <- r + r (1 - x r), each of which doubles the correct bits,
// so the result is within about 1 ulp of 1/x (an IEEE division's scaling and
// fix-up steps guard exponent extremes that the step's arguments, of order
// 1e-6 to 1e13, never reach).  Five instructions instead of about ten.
__device__ inline double drcp(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.);
    r = fma(r, e, r);
    e = fma(-x, r, 1.);
    return fma(r, e, r);
}
__host__ inline double drcp(double x) { return 1. / x; }

// FP64 elementary functions of the FP64 transfer (rays at the tau gate),
// written from their series for the arguments that transfer gives them
// (finite; exp's below 700 in magnitude, log's positive and normal):
//  dexp    e^x = 2^k e^r, k = rint(x / ln 2), r = x - k ln 2 with ln 2 in two
//          parts (each product exact inside its FMA), |r| <= ln 2 / 2; e^r by
//          its Taylor series to r^13/13! (truncation < 5e-18) in Horner form,
//          the last step fma(r, q, 1) rounded once, so that e^x - 1 for the
//          Planck factor's tiny x keeps the rounding of a correctly rounded exp
//  dexpm1  e^x - 1: the same series without its 1 for |x| <= 1/4
//  dlog    ln x = e ln 2 + 2 atanh s, x = 2^e m with m in [sqrt(1/2), sqrt 2),
//          s = (m - 1)/(m + 1), |s| <= 0.1716: atanh's series to s^25/25
//          (truncation < 2e-22)
//  dsqrt   x / sqrt(x) from the hardware reciprocal square root refined by
//          two Newton steps, then one correction of the product
// Each is a few dozen instructions, against a few hundred for the general
// library functions with their special cases.
constexpr double kLn2Hi = 0.6931471805599453, kLn2Lo = 2.3190468138462996e-17;

__host__ __device__ inline double exp_series_q(double r)   // (e^r - 1) / r
{
    double q = 1. / 6227020800.;
    q = fma(q, r, 1. / 479001600.);
    q = fma(q, r, 1. / 39916800.);
    q = fma(q, r, 1. / 3628800.);
    q = fma(q, r, 1. / 362880.);
    q = fma(q, r, 1. / 40320.);
    q = fma(q, r, 1. / 5040.);
    q = fma(q, r, 1. / 720.);
    q = fma(q, r, 1. / 120.);
    q = fma(q, r, 1. / 24.);
    q = fma(q, r, 1. / 6.);
    q = fma(q, r, 0.5);
    return fma(q, r, 1.);
}

__host__ __device__ inline double dexp(double x)
{
    const double k = rint(x * 1.4426950408889634);
    const double r = fma(-k, kLn2Lo, fma(-k, kLn2Hi, x));
    return ldexp(fma(r, exp_series_q(r), 1.), (int)k);
}

__host__ __device__ inline double dexpm1(double x)
{
    return fabs(x) <= 0.25 ? x * exp_series_q(x) : dexp(x) - 1.;
}

__host__ __device__ inline double dlog(double x)
{
    int e;
    double m = frexp(x, &e);
    if (m < 0.7071067811865476) {
        m *= 2.;
        e -= 1;
    }
    const double t = (m - 1.) * drcp(m + 1.);
    const double t2 = t * t;
    double q = 1. / 25.;
    for (int n = 23; n >= 3; n -= 2)
        q = fma(q, t2, 1. / n);
    q = fma(q, t2, 1.);
    const double fe = (double)e;
    return fma(fe, kLn2Hi, fma(fe, kLn2Lo, 2. * t * q));
}

__device__ inline double dsqrt(double x)
{
    double r = __builtin_amdgcn_rsq(x);
    r = r * fma(-0.5 * x * r, r, 1.5);
    r = r * fma(-0.5 * x * r, r, 1.5);
    double y = x * r;
    y = fma(0.5 * r, fma(-y, y, x), y);
    return x > 0. ? y : (x == 0. ? 0. : NAN);
}
__host__ inline double dsqrt(double x) { return sqrt(x); }

// theta = pi X2 + (1-h)/2 sin(2 pi X2) is evaluated as pi * (X2 + (1-h)/(2 pi)// theta = pi X2 + (1-h)/2 sin(2 pi X2) is evaluated as pi * (X2 + (1-h)/(2 pi)
// sin(2 pi X2)), so both sine/cosine pairs take arguments in units of pi and
// reduce exactly (sincospi) instead of through a general argument reduction.
__host__ __device__ inline Point make_point(double X1, double X2, double h)
{
    Point p;
    p.r = exp(X1);
    sincospi(2. * X2, &p.s2, &p.c2);
    const double tp = X2 + (1. - h) * (0.5 / kPi) * p.s2;
    sincospi(tp, &p.s, &p.c);
    return p;
}

// The RK2 midpoint and end point differ from the step's start point by
// e1 = dl k^1 / 2 (or dl k^1) in X1 and e2 = dl k^2 / 2 (or dl k^2) in X2,
// with |dl k^1| <= STEPSIZE and |dl k^2| <= STEPSIZE min(X2, 1-X2) by the
// step-size rule, so their trigonometry follows from the start point's by
// short Taylor series and angle addition.  With the reference's STEPSIZE =
// 0.03: |e1| <= 0.03, |2 pi e2| <= 0.0943, and the theta increment
// pi (e2 + (1-h)/(2 pi) (sin 2 pi X2' - sin 2 pi X2)) is at most
// pi |e2| + |2 pi e2| / 2 <= 0.0943.  Within the guards (0.031, 0.095) the
// truncation errors are below 2.2e-17; beyond them (only with a larger
// STEPSIZE) the point is evaluated in full.  Carried from step to step, the
// rounding of the additions accumulates as a random walk: a ray of 2000
// steps keeps its trigonometry within 1e-14 of a fresh evaluation.
__host__ __device__ inline double exp_small(double x)   // |x| <= 0.031: x^8/8! < 2.2e-17
{
    double p = 1. / 5040.;
    p = p * x + 1. / 720.;
    p = p * x + 1. / 120.;
    p = p * x + 1. / 24.;
    p = p * x + 1. / 6.;
    p = p * x + 0.5;
    p = p * x + 1.;
    return p * x + 1.;
}

// |x| <= 0.095: truncation x^11/11! < 1.5e-19 (sin), x^10/10! < 1.7e-17 (cos)
__host__ __device__ inline void sincos_small(double x, double &s, double &c)
{
    const double x2 = x * x;
    double ps = 1. / 362880.;
    ps = ps * x2 - 1. / 5040.;
    ps = ps * x2 + 1. / 120.;
    ps = ps * x2 - 1. / 6.;
    s = x + x * x2 * ps;
    double pc = -1. / 40320.;
    pc = pc * x2 + 1. / 720.;
    pc = pc * x2 - 1. / 24.;
    pc = pc * x2 + 0.5;
    c = 1. - x2 * pc;
}

// The series are evaluated on every lane and the guards tested once, at the
// end: the full evaluation replaces the result only beyond them.  UG (the
// kernels that trace rays from the camera) runs it in a wave-uniform branch,
// taken only when some lane needs it (approach kernel -6%, the lead
// kernel's long rays -9%); the resuming kernels, which run slower so, return
// it from a per-lane branch (one test instead of one per guard: their trace
// kernels -3 to -6%).
template <bool UG = false>
__host__ __device__ inline Point shifted_point(const Point &p0, double X1m, double X2m,
                                               double e1, double e2, double h)
{
    const double kh = (1. - h) * (0.5 / kPi);
    const double a2 = 2. * kPi * e2;
    Point p;
    p.r = p0.r * exp_small(e1);
    double ds, dc;
    sincos_small(a2, ds, dc);
    p.s2 = p0.s2 * dc + p0.c2 * ds;
    p.c2 = p0.c2 * dc - p0.s2 * ds;
    const double at = kPi * (e2 + kh * (p.s2 - p0.s2));   // theta_m - theta_0
    sincos_small(at, ds, dc);
    p.s = p0.s * dc + p0.c * ds;
    p.c = p0.c * dc - p0.s * ds;
    const bool ok = fabs(e1) <= 0.031 && fabs(a2) <= 0.095 && fabs(at) <= 0.095;
    if (UG) {
        if (__builtin_expect(wave_any(!ok), 0)) {
            const Point q = make_point(X1m, X2m, h);
            if (!ok)
                p = q;
        }
    } else if (__builtin_expect(!ok, 0)) {
        return make_point(X1m, X2m, h);
    }
    return p;
}

// Covariant metric (metric.c:94-138), 10 unique entries:
// g00 g01 g03 g11 g13 g22 g33 (g02 = g12 = g23 = 0).
struct MetricDD {
    double g00, g01, g03, g11, g13, g22, g33;
};

__host__ __device__ inline MetricDD metric_dd(const Point &p, double a, double h)
{
    MetricDD g;
    double r = p.r;
    double sin2 = p.s * p.s;
    double rho2 = r * r + a * a * p.c * p.c;
    double hfac = kPi + (1. - h) * kPi * p.c2;
    double z = 2. * r / rho2;
    g.g00 = -1. + z;
    g.g01 = z * r;
    g.g03 = -2. * a * r * sin2 / rho2;
    g.g11 = (1. + z) * r * r;
    g.g13 = -a * sin2 * (1. + z) * r;
    g.g22 = rho2 * hfac * hfac;
    g.g33 = sin2 * (rho2 + a * a * sin2 * (1. + z));
    return g;
}

__host__ __device__ inline void lower(const MetricDD &g, const double v[4], double o[4])
{
    o[0] = g.g00 * v[0] + g.g01 * v[1] + g.g03 * v[3];
    o[1] = g.g01 * v[0] + g.g11 * v[1] + g.g13 * v[3];
    o[2] = g.g22 * v[2];
    o[3] = g.g03 * v[0] + g.g13 * v[1] + g.g33 * v[3];
}

// Geodesic acceleration A^m = -Gamma^m_ab k^a k^b in MKS coordinates
// (integrator.c:126-146 with the MKS connection of metric.c:538-638), from
// the Kerr-Schild structure of the metric of metric.c:94-138: in KS
// coordinates (t, r, theta, phi) g = eta + z l l with z = 2r/Sigma,
// l = (1, 1, 0, -a sin^2), eta the flat metric in oblate coordinates, and
// g^-1 with 6 non-zero entries (g^tt = -(1+z), g^tr = z, g^rr = Delta/Sigma,
// g^rphi = a/Sigma, g^thth = 1/Sigma, g^phph = 1/(Sigma sin^2)).  With the KS
// wave vector K = (k^0, r k^1, theta' k^2, k^3) and P = l.K:
//   d_r g . K     = (z_r P, z_r P, 2r K^th, -a sin^2 z_r P + 2r sin^2 K^ph)
//   d_th g . K    = (z_th P - a z hs K^ph, z_th P - a (1+z) hs K^ph,
//                    -a^2 hs K^th,
//                    -a sin^2 z_th P - a hs (z K^t + (1+z) K^r) + E K^ph)
// with hs = d sin^2/d theta = 2 sin cos and E = hs (Sigma + a^2 sin^2 (1+2z));
// V_m = Gamma_mab K^a K^b = K^r (d_r g.K)_m + K^th (d_th g.K)_m
//       - 1/2 delta_m^r K.(d_r g.K) - 1/2 delta_m^th K.(d_th g.K),
// A_KS = -g^-1 V, and the MKS components follow from r = exp(X1) and
// theta(X2): A^1 = A^r / r - (k^1)^2, A^2 = (A^th - theta'' (k^2)^2) / theta'.
// The same mathematics as the reference's connection, with about 100
// operations instead of its 40 entries and 64-term contraction.
__host__ __device__ inline void geodesic_accel(const Point &p, double a, double h,
                                               const double k[4], double A[4])
{
    const double r = p.r, s = p.s, c = p.c;
    const double r2 = r * r, a2 = a * a;
    const double S2 = s * s;
    const double a2c2 = a2 * (c * c);
    const double Sig = r2 + a2c2;
    const double tr = 2. * r;
    const double Del = (r2 + a2) - tr;
    const double thp = kPi * (1. + (1. - h) * p.c2);          // d theta / d X2
    const double thpp = -2. * kPi * kPi * (1. - h) * p.s2;    // d^2 theta / d X2^2
    const double hs = 2. * s * c;
    // 1/(Sigma sin^2), 1/Sigma, 1/r and 1/theta' from one reciprocal
    const double SS2 = Sig * S2, rth = r * thp;
    const double inv = drcp(SS2 * rth);
    const double iSS2 = inv * rth;
    const double iS = iSS2 * S2;
    const double irth = inv * SS2;
    const double ir = irth * thp;
    const double ithp = irth * r;
    const double z = tr * iS;
    const double zr = 2. * (a2c2 - r2) * (iS * iS);           // dz/dr
    const double zt = z * a2 * hs * iS;                       // dz/dtheta
    const double Kt = k[0], Kr = r * k[1], Kh = thp * k[2], Kp = k[3];
    const double aS2 = a * S2;
    const double Ktr = Kt + Kr;
    const double Pl = Ktr - aS2 * Kp;
    const double zrP = zr * Pl, ztP = zt * Pl;
    const double Kh2 = Kh * Kh, Kp2 = Kp * Kp;
    const double ahs = a * hs;
    const double azhs = z * ahs;
    const double zKtr = z * Ktr + Kr;                         // z K^t + (1 + z) K^r
    const double E = hs * (Sig + a2 * S2 * (1. + 2. * z));
    const double Vt = Kr * zrP + Kh * (ztP - azhs * Kp);
    const double Vr = Kr * zrP + Kh * (ztP - (ahs + azhs) * Kp) - 0.5 * zrP * Pl -
                      r * (Kh2 + S2 * Kp2);
    const double KDh = ztP * Pl - 2. * ahs * Kp * zKtr - a2 * hs * Kh2 + E * Kp2;
    const double Vh = tr * Kr * Kh - a2 * hs * Kh2 - 0.5 * KDh;
    const double Vp = Kr * (tr * S2 * Kp - aS2 * zrP) + Kh * (E * Kp - aS2 * ztP - ahs * zKtr);
    const double At = (1. + z) * Vt - z * Vr;
    const double Ar = -(z * Vt + Del * iS * Vr + a * iS * Vp);
    const double Ah = -Vh * iS;
    const double Ap = -(a * iS * Vr + iSS2 * Vp);
    A[0] = At;
    A[1] = Ar * ir - k[1] * k[1];
    A[2] = (Ah - thpp * k[2] * k[2]) * ithp;
    A[3] = Ap;
}

// integrator.c:109-122: dl = -max(1 / sum_i 1/dl_i, 1e-12) with
// dl_1 = eps/|k^1|, dl_2 = eps min(X2, 1-X2)/|k^2|, dl_3 = eps/|k^3|, written
// as eps m / (m (|k^1| + |k^3|) + |k^2|) with m = min(X2, 1-X2): one
// reciprocal instead of four divisions.  The reference's 1e-40 guards change
// the result only when a component or m is exactly 0; then this form gives
// 0 or NaN (0 * inf), and fmax returns the 1e-12 floor as the reference does.
__host__ __device__ inline double stepsize(const double y[8], double eps)
{
    const double m = fabs(fmin(y[2], 1. - y[2]));
    const double den = m * (fabs(y[5]) + fabs(y[7])) + fabs(y[6]);
    return -fmax(eps * m * drcp(den), 1e-12);
}

// One RK2 midpoint step (integrator.c:44-70).  `pt` holds the trigonometry of
// y's point on entry and of the new point on exit.  The time coordinate
// X^0 is not advanced: nothing on the path reads it (the metric is
// stationary; the reference only prints it in its debug geodesic dump).
template <bool UG = false>
__host__ __device__ inline void rk2_step_from(double y[8], Point &pt, double dl, double a,
                                              double h, const double A0[4])
{
    double A[4], ym[8];
    const double hdl = 0.5 * dl;   // exact; one FMA per component
    ym[0] = y[0];
    for (int i = 1; i < 4; i++)
        ym[i] = y[i] + hdl * y[4 + i];
    for (int i = 0; i < 4; i++)
        ym[4 + i] = y[4 + i] + hdl * A0[i];
    Point pm = shifted_point<UG>(pt, ym[1], ym[2], ym[1] - y[1], ym[2] - y[2], h);
    geodesic_accel(pm, a, h, ym + 4, A);
    const double X1o = y[1], X2o = y[2];
    for (int i = 1; i < 4; i++)
        y[i] = y[i] + dl * ym[4 + i];
    for (int i = 0; i < 4; i++)
        y[4 + i] = y[4 + i] + dl * A[i];
    pt = shifted_point<UG>(pt, y[1], y[2], y[1] - X1o, y[2] - X2o, h);
}

// rk2_step_from with the start point's acceleration A0 given: the same
// arithmetic, so a caller that has A0 already (computed beside the previous
// step's transfer sample) gets the same bits.
template <bool UG = false>
__host__ __device__ inline void rk2_step(double y[8], Point &pt, double dl, double a, double h)
{
    double A0[4];
    geodesic_accel(pt, a, h, y + 4, A0);
    rk2_step_from<UG>(y, pt, dl, a, h, A0);
}

// The steppers of integrator.c:258-263, chosen by int_method
// (parameters.h:186-189); RK2 is the reference's default.
constexpr int kRK2 = 0, kRK4 = 1, kVerlet = 2;

// integrator.c:13-39: classical RK4, stages at y, y + dx0/2, y + dx1/2 and
// y + dx2 with dx_q = dl f(stage q), update y + (dx0 + 2 dx1 + 2 dx2 + dx3)/6,
// summed in that order.  Each stage point's trigonometry follows from the
// step's start point by angle addition (full evaluation beyond the guards).
__host__ __device__ inline void rk4_step(double y[8], Point &pt, double dl, double a, double h)
{
    const double w[3] = {0.5, 0.5, 1.};
    double ys[8], S[8], A[4];
    for (int i = 0; i < 8; i++)
        ys[i] = y[i];
    Point ps = pt;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        geodesic_accel(ps, a, h, ys + 4, A);
        double dx[8];
        for (int i = 1; i < 4; i++)
            dx[i] = dl * ys[4 + i];
        for (int i = 0; i < 4; i++)
            dx[4 + i] = dl * A[i];
        const double c = (q == 1 || q == 2) ? 2. : 1.;
        for (int i = 1; i < 8; i++)
            S[i] = q == 0 ? dx[i] : S[i] + dx[i] * c;
        if (q < 3) {
            for (int i = 1; i < 8; i++)
                ys[i] = y[i] + dx[i] * w[q];
            ps = shifted_point(pt, ys[1], ys[2], ys[1] - y[1], ys[2] - y[2], h);
        }
    }
    const double X1o = y[1], X2o = y[2];
    for (int i = 1; i < 8; i++)
        y[i] = y[i] + 1. / 6. * S[i];
    pt = shifted_point(pt, y[1], y[2], y[1] - X1o, y[2] - X2o, h);
}

// integrator.c:71-105: velocity Verlet (Dolence et al. 2009, eq. 14a-d):
// X' = X + dl k + dl^2 A(X, k)/2, k~ = k + dl A(X, k),
// k' = k + dl (A(X, k) + A(X', k~))/2.  The second acceleration is at the
// new point, whose trigonometry the next step reuses.
__host__ __device__ inline void verlet_step(double y[8], Point &pt, double dl, double a,
                                            double h)
{
    double A0[4], A1[4], ys[8];
    geodesic_accel(pt, a, h, y + 4, A0);
    ys[0] = y[0];
    for (int i = 1; i < 4; i++)
        ys[i] = y[i] + (dl * y[4 + i] + 0.5 * dl * dl * A0[i]);
    for (int i = 0; i < 4; i++)
        ys[4 + i] = y[4 + i] + A0[i] * dl;
    const Point pn = shifted_point(pt, ys[1], ys[2], ys[1] - y[1], ys[2] - y[2], h);
    geodesic_accel(pn, a, h, ys + 4, A1);
    for (int i = 1; i < 4; i++)
        y[i] = ys[i];
    for (int i = 0; i < 4; i++)
        y[4 + i] += 0.5 * (A0[i] + A1[i]) * dl;
    pt = pn;
}

template <int INT, bool UG = false>
__host__ __device__ inline void geodesic_step(double y[8], Point &pt, double dl, double a,
                                              double h)
{
    if (INT == kRK4)
        rk4_step(y, pt, dl, a, h);
    else if (INT == kVerlet)
        verlet_step(y, pt, dl, a, h);
    else
        rk2_step<UG>(y, pt, dl, a, h);
}

// Pixel part of initialize_photon (metric.c:650-727): k_mu from impact
// parameters, raised with the BL metric, converted to KS and to MKS X2.
__host__ __device__ inline void init_photon(const Camera &cm, double alpha, double beta,
                                            double y[8])
{
    double ll = -alpha * cm.sin_i;
    double qq = beta * beta + cm.mu0sq * (alpha * alpha - 1.);
    double sgn = (beta < 0.) ? -1. : ((beta > 0.) ? 1. : 0.);
    double kd2 = sgn * sqrt(fabs(qq - ll * ll * cm.cot2 + cm.cosc2));
    double ku0 = cm.g00 * (-1.) + cm.g03 * ll;
    double ku3 = cm.g33 * ll + cm.g03 * (-1.);
    double ku2 = cm.g22 * kd2;
    double ku1 = sqrt((-ku0 * (-1.) - ku2 * kd2 - ku3 * ll) / cm.g11);
    y[0] = 0.;
    y[1] = cm.X1;
    y[2] = cm.X2;
    y[3] = cm.X3;
    y[4] = ku0 + cm.t01 * ku1;
    y[5] = ku1;
    y[6] = ku2 * cm.thp_inv;
    y[7] = ku3 + cm.t31 * ku1;
}

// Interpolated primitives at a point (grmhd.c:35-71, with the 3D extension).
// The cell index and the weights come from the FP64 position; the weighted
// sums are formed in T (FP32 for the transfer pass, FP64 for the re-trace of
// rays near the tau gate) from the FP32 primitives.
// prims: 8 floats per cell, cells ordered ((i*n2 + j)*n3 + k).
template <bool DIM3, class T = float>
__host__ __device__ inline void sample_prims(const Params &P, const float4 *__restrict__ prims,
                                             const double y[8], T v[8])
{
    // grmhd.c:46-71: i = (int)(t - 1/2) truncates toward zero, the clamps set
    // del to 0 / 1, and otherwise del = t - (i + 1/2) with t = (X - x0)/dx
    const double t1 = (y[1] - P.startx1) * P.idx1;
    const double t2 = (y[2] - P.startx2) * P.idx2;
    int i = (int)(t1 - 0.5);
    int j = (int)(t2 - 0.5);
    double d1, d2;
    if (i < 0) {
        i = 0;
        d1 = 0.;
    } else if (i > P.n1 - 2) {
        i = P.n1 - 2;
        d1 = 1.;
    } else {
        d1 = t1 - (i + 0.5);
    }
    if (j < 0) {
        j = 0;
        d2 = 0.;
    } else if (j > P.n2 - 2) {
        j = P.n2 - 2;
        d2 = 1.;
    } else {
        d2 = t2 - (j + 0.5);
    }
    const T f1 = (T)d1, f2 = (T)d2;
    const T c0 = (T(1) - f1) * (T(1) - f2), c1 = (T(1) - f1) * f2;
    const T c2 = f1 * (T(1) - f2), c3 = f1 * f2;
#define RT_BI(f00, f01, f10, f11) ((f00) * c0 + (f01) * c1 + (f10) * c2 + (f11) * c3)
    // float4 element e of the staged primitives: a 32-bit byte offset from
    // the (wave-uniform) base, so that loads take the base from scalar
    // registers and the offset from one vector register, and the cell
    // arithmetic stays 32-bit (kMaxCells bounds it)
    const char *const pb = (const char *)prims;
    auto ld = [pb](unsigned e) { return *(const float4 *)(pb + (unsigned)(e * 16u)); };
    const unsigned b00 = (unsigned)i * P.s1 + (unsigned)j * P.s2;
    if (!DIM3) {
        const unsigned b10 = b00 + P.s1;
        float4 p00a = ld(b00), p00b = ld(b00 + 1);
        float4 p01a = ld(b00 + 2), p01b = ld(b00 + 3);
        float4 p10a = ld(b10), p10b = ld(b10 + 1);
        float4 p11a = ld(b10 + 2), p11b = ld(b10 + 3);
        v[0] = RT_BI(p00a.x, p01a.x, p10a.x, p11a.x);
        v[1] = RT_BI(p00a.y, p01a.y, p10a.y, p11a.y);
        v[2] = RT_BI(p00a.z, p01a.z, p10a.z, p11a.z);
        v[3] = RT_BI(p00a.w, p01a.w, p10a.w, p11a.w);
        v[4] = RT_BI(p00b.x, p01b.x, p10b.x, p11b.x);
        v[5] = RT_BI(p00b.y, p01b.y, p10b.y, p11b.y);
        v[6] = RT_BI(p00b.z, p01b.z, p10b.z, p11b.z);
        v[7] = RT_BI(p00b.w, p01b.w, p10b.w, p11b.w);
    } else {
        // periodic in X3 = phi; cell centres at startx3 + (k + 1/2) dx3.  The
        // remainder x - P floor(x / P) (one FMA) is fmod's up to a rounding
        // when x / P rounds across an integer, which the fix-up brings back
        // into [0, P); the weights are continuous there, so either side gives
        // the same sample to rounding.
        const double x3 = y[3] - P.startx3;
        double ph = fma(-floor(x3 * P.iperiod3), P.period3, x3);
        if (ph < 0.)
            ph += P.period3;
        else if (ph >= P.period3)
            ph -= P.period3;
        double t = ph * P.idx3 - 0.5;
        double fk = floor(t);
        const T d3 = (T)(t - fk);
        // t lies in [-1/2, n3 - 1/2), so k in [-1, n3 - 1] (clamped: a NaN
        // position converts to an arbitrary integer on some targets)
        int k = (int)fk;
        k = k < -1 ? -1 : (k > P.n3 - 1 ? P.n3 - 1 : k);
        const int k0 = k < 0 ? P.n3 - 1 : k, k1 = k + 1 == P.n3 ? 0 : k + 1;
        const unsigned b01 = b00 + P.s2, b10 = b00 + P.s1, b11 = b10 + P.s2;
        T w[2][8];
        const int ks[2] = {k0, k1};
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const unsigned kk = 2u * (unsigned)ks[q];
            float4 p00a = ld(b00 + kk), p00b = ld(b00 + kk + 1);
            float4 p01a = ld(b01 + kk), p01b = ld(b01 + kk + 1);
            float4 p10a = ld(b10 + kk), p10b = ld(b10 + kk + 1);
            float4 p11a = ld(b11 + kk), p11b = ld(b11 + kk + 1);
            w[q][0] = RT_BI(p00a.x, p01a.x, p10a.x, p11a.x);
            w[q][1] = RT_BI(p00a.y, p01a.y, p10a.y, p11a.y);
            w[q][2] = RT_BI(p00a.z, p01a.z, p10a.z, p11a.z);
            w[q][3] = RT_BI(p00a.w, p01a.w, p10a.w, p11a.w);
            w[q][4] = RT_BI(p00b.x, p01b.x, p10b.x, p11b.x);
            w[q][5] = RT_BI(p00b.y, p01b.y, p10b.y, p11b.y);
            w[q][6] = RT_BI(p00b.z, p01b.z, p10b.z, p11b.z);
            w[q][7] = RT_BI(p00b.w, p01b.w, p10b.w, p11b.w);
        }
        for (int q = 0; q < 8; q++)
            v[q] = (T(1) - d3) * w[0][q] + d3 * w[1][q];
    }
#undef RT_BI
}

// The fluid state the transfer needs at one sample (FP32).
struct Fluid {
    float ne, thetae, B;   // cgs density, dimensionless temperature, gauss
    float mu;              // cos(pitch angle)
    float gdop;            // nu_plasma / nu_observed = -(u . k)
    float dthetae;         // d Thetae / d R_high
};

// Covariant MKS metric in FP32 from the point's FP64 trigonometry.
struct MetricF {
    float g00, g01, g03, g11, g13, g22, g33;
    float gcon00, gcon01;  // the two inverse-metric entries the fluid needs
};

__host__ __device__ inline MetricF metric_f32(const Point &pt, const Params &P)
{
    MetricF g;
    const float r = (float)pt.r, s = (float)pt.s, c = (float)pt.c;
    const float a = P.fa;
    const float sin2 = s * s;
    const float rho2 = r * r + a * a * c * c;
    const float hfac = P.fpi_h * (float)pt.c2 + 3.14159265f;
    const float irho2 = frcp(rho2);
    const float z = 2.f * r * irho2;
    g.g00 = -1.f + z;
    g.g01 = z * r;
    g.g03 = -a * z * sin2;
    g.g11 = (1.f + z) * r * r;
    g.g13 = -a * sin2 * (1.f + z) * r;
    g.g22 = rho2 * hfac * hfac;
    g.g33 = sin2 * (rho2 + a * a * sin2 * (1.f + z));
    g.gcon00 = -1.f - z;
    g.gcon01 = 2.f * irho2;
    return g;
}

__host__ __device__ inline void lower_f(const MetricF &g, const float v[4], float o[4])
{
    o[0] = g.g00 * v[0] + g.g01 * v[1] + g.g03 * v[3];
    o[1] = g.g01 * v[0] + g.g11 * v[1] + g.g13 * v[3];
    o[2] = g.g22 * v[2];
    o[3] = g.g03 * v[0] + g.g13 * v[1] + g.g33 * v[3];
}

// get_fluid_params (raptor_harm_model.c:190-320) + pitch_angle and the
// plasma-frame frequency factor (radiative_transfer.c:17-47), in FP32.
// Returns false when b^2/rho > 1 (the sample is skipped).  The electron
// temperature ratio is formed as R_high + (R_low - R_high)/(1 + beta^2),
// which equals the reference's expression and stays finite when beta^2
// overflows FP32.
template <bool DIM3, bool NOEXIT = false>
__host__ __device__ inline bool fluid_sample(const Params &P, const float4 *__restrict__ prims,
                                             const Point &pt, const double y[8], Fluid &F,
                                             float *ucon_out = nullptr,
                                             float *bcon_out = nullptr,
                                             float *beta_bern_out = nullptr)
{
    float v[8];
    sample_prims<DIM3>(P, prims, y, v);
    const float rho = v[0], uu = v[1];
    const MetricF g = metric_f32(pt, P);
    const float V1 = v[2], V2 = v[3], V3 = v[4];
    float VdotV = g.g11 * V1 * V1 + g.g13 * V1 * V3 + g.g22 * V2 * V2 + g.g13 * V3 * V1 +
                  g.g33 * V3 * V3;
    float Vfac = fsqrt(-frcp(g.gcon00) * (1.f + fabsf(VdotV)));
    float u[4], ud[4], b[4], bd[4];
    u[0] = -Vfac * g.gcon00;
    u[1] = V1 - Vfac * g.gcon01;
    u[2] = V2;
    u[3] = V3;
    lower_f(g, u, ud);
    float UdotBp = ud[1] * v[5] + ud[2] * v[6] + ud[3] * v[7];
    const float iu0 = frcp(u[0]);
    b[0] = UdotBp;
    b[1] = (v[5] + u[1] * UdotBp) * iu0;
    b[2] = (v[6] + u[2] * UdotBp) * iu0;
    b[3] = (v[7] + u[3] * UdotBp) * iu0;
    lower_f(g, b, bd);
    float Bsq = b[0] * bd[0] + b[1] * bd[1] + b[2] * bd[2] + b[3] * bd[3];
    F.ne = rho * P.fNe_unit;
    const float bnorm = fsqrt(Bsq);
    F.B = bnorm * P.fB_unit + 1e-20f;
    float beta = uu * (2.f / 3.f) * frcp(Bsq + 1e-20f);
    float b2 = beta * beta;
    float w = frcp(1.f + b2);
    float trat = P.fR_high + (P.fR_low - P.fR_high) * w;
    float itr2 = frcp(trat + 2.f);
    float gam2 = 0.5f * ((1.f + (2.f / 3.f) * (trat + 1.f) * itr2) + 4.f / 3.f);
    const float mr = (float)(kProtonMass / kElectronMass);
    const float itr1 = frcp(1.f + trat);
    const float irho = frcp(rho);
    const float uor = uu * irho;
    F.thetae = uor * ((gam2 - 1.f) * mr * itr1);
    float dgam2 = (1.f / 3.f) * itr2 * itr2;
    const float dtrat = (b2 <= 3e38f) ? b2 * w : 1.f;   // d trat / d R_high
    F.dthetae = uor * mr * (dgam2 * (1.f + trat) - (gam2 - 1.f)) * itr1 * itr1 * dtrat;
    if (F.thetae < 0.f) {
        F.thetae = 1e-20f;
        F.dthetae = 0.f;
    }
    if (ucon_out) {
        for (int q = 0; q < 4; q++) {
            ucon_out[q] = u[q];
            bcon_out[q] = b[q];
        }
    }
    if (beta_bern_out) {
        // raptor_harm_model.c:263, 281 (Bernoulli parameter with gamma = 4/3)
        beta_bern_out[0] = beta;
        beta_bern_out[1] = -(1.f + uor * (4.f / 3.f)) * ud[0];
    }
    // NOEXIT (the resuming kernels, 3% faster so; the kernels from the
    // camera run slower): the sigma cut's decision is returned at the end,
    // the pitch angle and Doppler factor computed on every lane (unused
    // where it cuts), so that the sample has no divergent exit
    const bool cut = Bsq * irho > 1.f;
    if (!NOEXIT && cut)
        return false;
    // pitch angle (cos) and Doppler factor with k lowered by the same metric
    float k[4] = {(float)y[4], (float)y[5], (float)y[6], (float)y[7]};
    float kd[4];
    lower_f(g, k, kd);
    float ku = kd[0] * u[0] + kd[1] * u[1] + kd[2] * u[2] + kd[3] * u[3];
    float mu = (kd[0] * b[0] + kd[1] * b[1] + kd[2] * b[2] + kd[3] * b[3]) *
               frcp(fabsf(ku) * (F.B * P.fiB_unit));
    // radiative_transfer.c:38-42 clamps with mu / |mu|, which is exactly +-1
    if (fabsf(mu) > 1.f)
        mu = copysignf(1.f, mu);
    F.mu = mu;
    F.gdop = -ku;
    return !cut;
}

// Thermal synchrotron emissivity (j_nu.c:130-153) and Kirchhoff absorptivity
// (j_nu.c:188-200) in FP32 and scaled units:
//   jh   = j_nu / (e^2/c) without its exp(-X^(1/3)) factor; kept as
//          n_e sqrt(2) pi / 3 * (nu_s / K2) f with nu_s / K2 = nu_c sin / 9
//          since K2 = 2 Theta_e^2 cancels Theta_e^2 in nu_s;
//   cb   = X^(1/3), so j_nu = (e^2/c) jh exp(-cb);
//   kap  = alpha_nu / (jh exp(-cb)) = e2_ch (exp(x)-1) c^2 / (2 nu_p^3),
//          x = h nu_p / (k T_e), the reference's Planck function without h.
// jh == 0 marks the reference's zero-emissivity gates.  When j is zero and the
// reference's FP64 exp(x) overflows (x > 709.78, only for Theta_e below
// ~1e-11) its absorptivity is 0/0 = NaN; `poison` reports that case.
// The fit is split into what one fluid sample fixes (EmisPre, once per
// sample) and what each frequency adds (emission_nu).  X^(1/2) and X^(1/3)
// are powers of s6 = X^(1/6), one exp2 for all three roots.  The caller
// passes s6 and kf = (e^2 / (c h)) c^2 / (2 nu_p^3): with nu_p = g nu, the
// transfer forms them from per-sample and per-frequency factors,
// log2 s6 = (log2(g 2^37 / nu_s) + log2(nu / 2^37)) / 6 and kf = g^-3 kf(nu);
// both logs stay of the size of log2 X near 2^37 Hz = 137 GHz, so their sum
// loses no more than log2 X itself would.
struct EmisF {
    float jh, cb, kap, s6;
    bool poison;
};

struct EmisPre {
    float inus;   // 1 / nu_s
    float jpre;   // n_e sqrt(2) pi / 27 nu_c sin
    float xk;     // h / (m_e c^2 Theta_e): x = xk nu_p
    float Te;
    bool on;      // the sin and Theta_e gates pass
};

__host__ __device__ inline EmisPre emission_pre(float B, float mu, float Te, float ne)
{
    EmisPre p;
    const float sth = fsqrt(1.f - mu * mu);
    const float nu_c = (float)(kElectronCharge / (2. * kPi * kElectronMass * kC)) * B;
    const float nu_s = (2.f / 9.f) * nu_c * Te * Te * sth;
    p.inus = frcp(nu_s);
    p.jpre = ne * (float)(1.4142135623730951 * kPi / 3. / 9.) * (nu_c * sth);
    p.xk = (float)(kPlanck / (kElectronMass * kC * kC)) * frcp(Te);
    p.Te = Te;
    p.on = !(fabsf(sth) < 1.e-10f) && Te > 1e-4f;
    return p;
}

// The FP64 overflow test of the zero-emissivity case (poison) for a plasma
// frequency nu_p at dimensionless temperature Te.
__host__ __device__ inline bool planck_poison(float Te, float nu_p)
{
    const double T = (double)Te * kElectronMass * kC * kC / kBoltzmann;
    return kPlanck * (double)nu_p / (kBoltzmann * T) > 709.782712893384;
}

// POISON = false leaves e.poison false: for callers that apply the (rare)
// Theta_e <= 1e-4 case once per sample instead of once per frequency.
template <bool POISON = true>
__host__ __device__ inline EmisF emission_nu(const EmisPre &p, float nu_p, float s6, float kf)
{
    EmisF e;
    e.poison = false;
    if constexpr (!POISON) {
        // branch-free: the fit's factors for every lane, zeros selected
        // where the gates close (the lanes' values are the same as below;
        // a gated lane's factors may be inf or NaN and are discarded)
        const bool zero = nu_p == 1.f || !p.on;
        const float cb = s6 * s6;
        const float ff = cb * s6 + 1.8877486253633868f * s6;
        const float kap = fexpm1(p.xk * nu_p) * kf;
        e.cb = zero ? 0.f : cb;
        e.s6 = zero ? 0.f : s6;
        e.jh = zero ? 0.f : p.jpre * (ff * ff);
        e.kap = zero ? 0.f : kap;
        return e;
    }
    if (nu_p == 1.f || !p.on) {
        e.jh = 0.f;
        e.cb = 0.f;
        e.kap = 0.f;
        e.s6 = 0.f;
        if (POISON && !(p.Te > 1e-4f))
            e.poison = planck_poison(p.Te, nu_p);
        return e;
    }
    const float cb = s6 * s6;                                         // X^(1/3)
    const float ff = cb * s6 + 1.8877486253633868f * s6;
    e.cb = cb;
    e.s6 = s6;
    e.jh = p.jpre * (ff * ff);
    e.kap = fexpm1(p.xk * nu_p) * kf;
    return e;
}

__host__ __device__ inline EmisF emission_f32(float B, float mu, float Te, float nu_p, float ne)
{
    const EmisPre p = emission_pre(B, mu, Te, ne);
    const float inu = frcp(nu_p), cn = (float)kC * inu;
    return emission_nu(p, nu_p, fexp2(flog2(nu_p * p.inus) * (1.f / 6.f)),
                       (float)(0.5 * kE2CH) * (cn * cn) * inu);
}

// d ln j / d ln X of the fit at fixed n_e and nu_s, from s6 = X^(1/6)
__host__ __device__ inline float dlnj_dlnX(float s6)
{
    const float c1112 = 1.8877486253633868f;
    const float cb = s6 * s6, sx = cb * s6;
    return 2.f * (0.5f * sx + c1112 / 6.f * s6) * frcp(sx + c1112 * s6) - cb * (1.f / 3.f);
}

// Emissivity and absorptivity of the other emission models: j_nu = (e^2/c) jx
// and alpha_nu, both cgs; `poison` as for the thermal fit.
struct EmisX {
    float jx, alpha;
    bool poison;
};

// log2(2^p + 2^q) without overflow
__host__ __device__ inline float log2_add(float p, float q)
{
    return fmaxf(p, q) + flog2(1.f + fexp2(-fabsf(p - q)));
}

// Gauss series sum_n (a)_n (b)_n / ((c)_n n!) x^n for 0 <= x <= 1/2, FP32.
__host__ __device__ inline float hyp_series(float a, float b, float c, float x)
{
    float term = 1.f, sum = 1.f;
    for (int n = 0; n < 96; n++) {
        const float fn = (float)n;
        term *= (a + fn) * (b + fn) * frcp((c + fn) * (fn + 1.f)) * x;
        sum += term;
        if (fabsf(term) <= 3e-8f * fabsf(sum))
            break;
    }
    return sum;
}

// 2F1(a, b; c; z) of the kappa absorptivity, z = -kappa Theta_e < 0
// (j_nu.c:76-91): Pfaff's transformation for -1 <= z < 0, the 1/(1-z)
// continuation the reference writes out for z < -1.
__host__ __device__ inline float kappa_hyp(const KappaConst &K, float z)
{
    const float l1z = flog2(1.f - z);
    if (z >= -1.f)
        return fexp2(-K.ha * l1z) * hyp_series(K.ha, K.hc - K.hb, K.hc, z * frcp(z - 1.f));
    // the first term's 2F1(a, c-b; a-b+1; x) has c - b = a - b + 1 (-1/3 for
    // any kappa), so it is (1-x)^-a and the term is g1 (-z)^-a
    const float x = frcp(1.f - z);
    return fexp2(-K.ha * flog2(-z)) * K.g1 +
           fexp2(-K.hb * l1z) * K.g2 * hyp_series(K.hb, K.hc - K.ha, K.hb - K.ha + 1.f, x);
}

// log2 of kappa_hyp(K, -2^u): cubic Hermite interpolation of the table
// (value and derivative at both ends of the interval), the series outside it.
__host__ __device__ inline float kappa_lhyp(const KappaConst &K, float u)
{
    const float t = (u - K.tu0) * K.tidu;
    const float fi = floorf(t);
    if (K.tab && fi >= 0.f && fi < (float)(K.tn - 1)) {
        const int i = (int)fi;
        const float f = t - fi, h = 1.f / K.tidu;
        const float2 p0 = K.tab[i], p1 = K.tab[i + 1];
        const float f2 = f * f, g = 1.f - f;
        return p0.x * ((1.f + 2.f * f) * g * g) + (p0.y * h) * (f * g * g) +
               p1.x * (f2 * (3.f - 2.f * f)) - (p1.y * h) * (f2 * g);
    }
    return flog2(kappa_hyp(K, -fexp2(u)));
}

// emission_coeff_kappa_FIT and absorption_coeff_kappa_FIT (j_nu.c:28-106,
// Pandya et al. 2016), kappa = 3.5, width w = Theta_e, mu = cos(pitch).
// j = (e^2/c) n_e nu_c sin J_s, alpha = n_e e / (B sin) A_s, with the
// power laws in X = nu / nu_s, nu_s = nu_c sin (w kappa)^2, taken in log2.
__host__ __device__ inline EmisX emission_kappa(const KappaConst &K, float B, float mu, float Te,
                                                float nu_p, float ne)
{
    EmisX e{0.f, 0.f, false};
    if (nu_p == 1.f || fabsf(mu) == 1.f)
        return e;
    const float kap = 3.5f;
    const float sth = fsqrt(1.f - mu * mu);
    const float nuc = (float)(kElectronCharge / (2. * kPi * kElectronMass * kC)) * B;
    const float wk = Te * kap;
    const float lwk = flog2(wk);
    const float lX = flog2(nu_p) - flog2(nuc * sth) - 2.f * lwk;
    const float lJ = -log2_add(-K.xj * (lX * (1.f / 3.f) + K.lJlow),
                               -K.xj * (-0.5f * (kap - 2.f) * lX + K.lJhigh)) *
                     frcp(K.xj);
    e.jx = fexp2(lJ) * (ne * nuc * sth);
    const float z = -wk;
    if (fabsf(z) == 1.f)
        return e;   // j_nu.c:79: the absorptivity is 0 there
    const float lhyp = K.tab ? kappa_lhyp(K, lwk) : flog2(kappa_hyp(K, z));
    const float lAl = -(5.f / 3.f) * lX + K.lAlow - (16.f / 3.f - kap) * lwk + lhyp;
    const float lAh = -0.5f * (3.f + kap) * lX + K.lAhigh - 5.f * lwk;
    const float lA = -log2_add(-K.xa * lAl, -K.xa * lAh) * frcp(K.xa);
    e.alpha = fexp2(lA + flog2(ne * (float)kElectronCharge * frcp(B * sth)));
    return e;
}

// emission_coeff_THSYNCHAV (j_nu.c:157-167, Dexter 2009) times e^2/c, with
// Kirchhoff's absorptivity (j_nu.c:194-200) as for the thermal fit.
__host__ __device__ inline EmisX emission_thsynchav(float B, float Te, float nu_p, float ne)
{
    EmisX e{0.f, 0.f, false};
    const float lnuc = flog2((float)(3. * kElectronCharge / (4. * kPi * kElectronMass * kC)) * B);
    const float lTe = flog2(Te);
    const float lx = flog2(nu_p) - lnuc - 2.f * lTe;   // log2 x_M
    const float I = 4.0505f * fexp2(-lx * (1.f / 6.f)) *
                    (1.f + 0.4f * fexp2(-0.25f * lx) + 0.5316f * fexp2(-0.5f * lx)) *
                    fexp(-1.8899f * fexp2(lx * (1.f / 3.f)));
    // n_e nu_p I / (2 sqrt 3 Theta_e^2), in log2 so that Theta_e^2 cannot
    // underflow where I has
    e.jx = fexp2(flog2(I) + flog2(nu_p * ne) - 2.f * lTe) * (float)(0.5 / 1.7320508075688772);
    const float inu = frcp(nu_p);
    const float x = (float)(kPlanck / (kElectronMass * kC * kC)) * nu_p * frcp(Te);
    const float cn = (float)kC * inu;
    if (e.jx != 0.f) {
        // below 2^-53 the reference's exp(x) - 1 (j_nu.c:190) is exactly 0, its
        // B_nu infinite and its absorptivity 0 (nu_p = 1 Hz, the floor of
        // radiative_transfer.c:24-26, where this fit has no gate)
        e.alpha = x < 1.1102230246251565e-16f
                      ? 0.f
                      : e.jx * ((float)(0.5 * kE2CH) * fexpm1(x) * (cn * cn) * inu);
    } else {
        double T = (double)Te * kElectronMass * kC * kC / kBoltzmann;
        e.poison = kPlanck * (double)nu_p / (kBoltzmann * T) > 709.782712893384;
    }
    return e;
}

// Per-ray accumulators for NF frequencies: optical depth (FP32 with Kahan
// compensation: the ln 1000 gate compares it), I_nu nu^3 (a plain FP32 sum of
// positive terms, relative rounding error about sqrt(n) 2^-24 for n samples,
// at most n 2^-24), and the tangents (FP64: they difference nearly equal
// terms, d ln j - d tau).
template <int NF, bool GRAD>
struct Accum {
    float tau[NF], ctau[NF];
    float I[NF];
    float wrel;   // sum over samples in the gate window of dI / I (see kGateMargin)
    unsigned gmask;   // bit f: frequency f had a sample in the gate window
    double dtau[GRAD ? NF : 1][2];
    double dI[GRAD ? NF : 1][2];
    __host__ __device__ void clear()
    {
        wrel = 0.f;
        gmask = 0u;
        for (int f = 0; f < NF; f++) {
            tau[f] = ctau[f] = 0.f;
            I[f] = 0.f;
            if (GRAD) {
                dtau[f][0] = dtau[f][1] = 0.;
                dI[f][0] = dI[f][1] = 0.;
            }
        }
    }
};

// The tau gate (radiative_transfer.c:90-95) is a discontinuity: a sample adds
// its emission only while tau < ln 1000.  The FP32 tau differs from an FP64
// one by up to ~3e-4 where it reaches the gate (measured over a 1024^2 image:
// the FP32 fluid sample's rounding, accumulated to tau ~ 7), so at samples
// whose FP32 tau lies within kGateMargin of ln 1000 the two may decide
// differently.  Each such sample adds its dI / I to Accum::wrel (I after the
// sample, or as if it had been added); a ray whose wrel exceeds kGateFlip is
// traced again with the FP64 transfer (transfer_sample_f64), so that no
// pixel can differ from an FP64 transfer by more than kGateFlip through a
// decision at the gate.  About 1 ray in 3000 at 1024^2 and 230 GHz.
constexpr float kGateMargin = 1e-3f;
constexpr float kGateFlip = 5e-4f;

// The FP64 transfer's optical depth lives in the FP32 pair (tau, ctau) of its
// accumulator, which it does not otherwise use (bit pattern of one double).
__host__ __device__ inline double tau_f64(const float &lo, const float &hi)
{
    uint32_t a, b;
    memcpy(&a, &lo, 4);
    memcpy(&b, &hi, 4);
    const uint64_t u = ((uint64_t)b << 32) | a;
    double d;
    memcpy(&d, &u, 8);
    return d;
}
__host__ __device__ inline void set_tau_f64(float &lo, float &hi, double d)
{
    uint64_t u;
    memcpy(&u, &d, 8);
    const uint32_t a = (uint32_t)u, b = (uint32_t)(u >> 32);
    memcpy(&lo, &a, 4);
    memcpy(&hi, &b, 4);
}

__host__ __device__ inline void kahan_add(float &s, float &c, float x)
{
    float yk = x - c;
    float t = s + yk;
    c = (t - s) - yk;
    s = t;
}

// radiative_transfer.c:71-98 for one sample: end-point quadrature with the
// step's |dlambda|, tau gate 0 <= tau < ln 1000.  With g = nu_p / nu and
// dl = |dlambda| R_g:
//   dtau = nu_p alpha_nu dl / nu = g alpha_nu dl
//   dI nu^3 = (j_nu / nu_p^2) exp(-tau) (dl / nu) nu^3 = j_nu exp(-tau) dl / g^2
// wf (the second trace of exact_retrace): each frequency's share of the
// samples in the gate window, instead of their sum in acc.wrel.  The
// accumulators' frequency f is the render's frequency f0 + f.
// A sample of frequency f whose tau is in the gate window: its share of the
// intensity, dI / I with I after this sample or as if it had been added
// (0, a first sample at the gate, gives an infinite share, which flags the
// ray), into wf[f] (the second trace) or into the accumulator's sum and mask.
template <int NF, bool GRAD>
__host__ __device__ inline void gate_share(Accum<NF, GRAD> &acc, int f, float tau, float dI,
                                           float *wf)
{
    const float lnth = 6.907755278982137f;
    const float Iw = tau < lnth ? acc.I[f] : acc.I[f] + dI;
    const float share = dI * frcp(Iw);
    if (wf) {
        wf[f] += share;
    } else {
        acc.wrel += share;
        acc.gmask |= 1u << f;
    }
}

// The thermal fit's emissivity factor of frequency f0 + f at a sample (the
// part of transfer_sample that the gate-window pass below evaluates again).
__host__ __device__ inline float plasma_nu(const Params &P, float gdop, int f)
{
    const float nu_p = gdop * P.fnu[f];
    return nu_p <= 0.f ? 1.f : nu_p;
}

__host__ __device__ inline EmisF thermal_nu(const Params &P, const EmisPre &pre, float gdop,
                                            float lgs, float rg3, int f)
{
    return emission_nu<false>(pre, plasma_nu(P, gdop, f),
                              fexp2(fmaf(lgs, 1.f / 6.f, P.flnu6[f])), rg3 * P.fkap[f]);
}

template <int NF, bool GRAD, int EMIS = kThSynch>
__host__ __device__ inline void transfer_sample(const Params &P, const Fluid &F, double dlabs,
                                                Accum<NF, GRAD> &acc, float *wf = nullptr,
                                                int f0 = 0)
{
    const float lnth = 6.907755278982137f;   // log(1000.)
    const float dl = (float)(dlabs * P.Rg);
    // For the thermal fit, g enters only where j != 0, hence nu_p > 0 and
    // g = gdop, so its reciprocals are taken once for all frequencies (with
    // j = 0 the factor only has to be finite: gdop <= 0 gives 0).
    const float rg = frcp(F.gdop);
    const float dlg2 = F.gdop > 0.f ? dl * frcp(F.gdop * F.gdop) : 0.f;
    const EmisPre pre = emission_pre(F.B, F.mu, F.thetae, F.ne);
    // log2 X at 2^37 Hz (the power of two scales exactly), to which each
    // frequency adds log2(nu / 2^37) in an explicit FMA: a separate product
    // would be contracted into the sum or not depending on how many
    // frequencies share it, and launches of different frequency counts must
    // give identical bits
    const float lgs = flog2(F.gdop * pre.inus * 0x1p37f);
    const float rg3 = rg * rg * rg;
    // Several frequencies: the smallest |tau - ln 1000| over them after this
    // sample, so that the gate window is tested once per sample, not once
    // per frequency (a branch per frequency held an exec mask each, and at 8
    // frequencies the trace kernel spilled scalar registers into vector lanes
    // and read them back every frequency).  One frequency: tested where its
    // dI is at hand (keeping the sample's emission factors for a second pass
    // would hold their registers through the transfer).
    [[maybe_unused]] float dev = INFINITY;
#pragma unroll
    for (int f = 0; f < NF; f++) {
        if constexpr (EMIS != kThSynch) {
            static_assert(!GRAD, "the tangents are those of the thermal fit");
            if (f >= P.nf)
                continue;   // (not break: the loop stays fully unrolled)
            float g = F.gdop;
            float nu_p = g * P.fnu[f0 + f];
            if (nu_p <= 0.f) {
                nu_p = 1.f;
                g = frcp(P.fnu[f0 + f]);
            }
            const EmisX e = EMIS == kKappa
                                ? emission_kappa(P.kc, F.B, F.mu, F.thetae, nu_p, F.ne)
                                : emission_thsynchav(F.B, F.thetae, nu_p, F.ne);
            const float dtau = (EMIS == kKappa || e.jx != 0.f) ? g * dl * e.alpha
                                                               : (e.poison ? NAN : 0.f);
            kahan_add(acc.tau[f], acc.ctau[f], dtau);
            const float tau = acc.tau[f];
            const float dI = (float)kE2C * e.jx * fexp(-tau) * (dl * frcp(g * g));
            if (tau < lnth && !(tau < 0.f))
                acc.I[f] += dI;
            continue;
        }
        const float g = F.gdop;
        const EmisF e = thermal_nu(P, pre, g, lgs, rg3, f0 + f);
        const float jx = e.jh * fexp(-e.cb);
        float dtau = e.jh != 0.f ? g * dl * jx * e.kap : 0.f;   // poison: after the loop
        kahan_add(acc.tau[f], acc.ctau[f], dtau);
        const float tau = acc.tau[f];
        const float dI = (float)kE2C * jx * fexp(-tau) * dlg2;
        float dlnj0 = 0.f, dlnj1 = 0.f, dlna1 = 0.f;
        if (GRAD && e.jh != 0.f) {
            float nu_p = g * P.fnu[f0 + f];
            float Dx = dlnj_dlnX(e.s6);
            float dlnB = 0.5f * (F.B - 1e-20f) * frcp(F.B);
            float dlnT = F.dthetae * frcp(F.thetae);
            float x = pre.xk * nu_p;
            float dlnBnu = x * frcp(-fexpm1(-x));
            dlnj0 = 1.f + dlnB * (1.f - Dx);
            dlnj1 = -2.f * Dx * dlnT;
            dlna1 = dlnj1 - dlnBnu * dlnT;
        }
        if (GRAD && dtau != 0.f) {
            acc.dtau[f][0] += (double)dtau * dlnj0;
            acc.dtau[f][1] += (double)dtau * dlna1;
        }
        if (tau < lnth && !(tau < 0.f)) {
            acc.I[f] += dI;
            if (GRAD && dI != 0.f) {
                acc.dI[f][0] += (double)dI * (dlnj0 - acc.dtau[f][0]);
                acc.dI[f][1] += (double)dI * (dlnj1 - acc.dtau[f][1]);
            }
        }
        if constexpr (NF == 1) {
            if (fabsf(tau - lnth) < kGateMargin && dI > 0.f)
                gate_share(acc, f, tau, dI, wf);
        } else {
            dev = fminf(dev, fabsf(tau - lnth));
        }
    }
    if constexpr (EMIS == kThSynch) {
        // Theta_e <= 1e-4: j = 0 at every frequency, and where the reference's
        // FP64 Planck exponent overflows its absorptivity is 0/0, which makes
        // tau NaN (and with it the rest of the ray's contributions).  The
        // loop above added dtau = 0 instead (and dI = 0, which changes no
        // sum); setting the accumulators NaN here gives the same state as
        // adding the NaN there, without a branch per frequency.
        if (__builtin_expect(!(pre.Te > 1e-4f), 0)) {
#pragma unroll
            for (int f = 0; f < NF; f++) {
                if (planck_poison(pre.Te, plasma_nu(P, F.gdop, f0 + f))) {
                    acc.tau[f] = acc.ctau[f] = NAN;
                    if (GRAD)
                        acc.dtau[f][0] = acc.dtau[f][1] = NAN;
                }
            }
        }
    }
    if constexpr (EMIS == kThSynch && NF > 1) {
        if (__builtin_expect(dev < kGateMargin, 0)) {
            // a frequency's tau is in the gate window: its dI again (the same
            // arithmetic), and its share of the intensity
#pragma unroll
            for (int f = 0; f < NF; f++) {
                const float tau = acc.tau[f];
                if (!(fabsf(tau - lnth) < kGateMargin))
                    continue;
                const EmisF e = thermal_nu(P, pre, F.gdop, lgs, rg3, f0 + f);
                const float jx = e.jh * fexp(-e.cb);
                const float dI = (float)kE2C * jx * fexp(-tau) * dlg2;
                if (dI > 0.f)
                    gate_share(acc, f, tau, dI, wf);
            }
        }
    }
}

// ---- FP64 transfer for rays flagged at the tau gate -------------------------
// get_fluid_params (raptor_harm_model.c:190-320), pitch_angle and
// freq_in_plasma_frame (radiative_transfer.c:17-47) in FP64, in the
// reference's form (its electron temperature ratio, metric and lowering),
// from the same FP32 primitives interpolated in FP64.
struct Fluid64 {
    double ne, thetae, B, mu, gdop, dthetae;
};

template <bool DIM3>
__host__ __device__ inline bool fluid_sample_f64(const Params &P, const float4 *__restrict__ prims,
                                                 const Point &pt, const double y[8], Fluid64 &F)
{
    double v[8];
    sample_prims<DIM3, double>(P, prims, y, v);
    const double rho = v[0], uu = v[1];
    const MetricDD g = metric_dd(pt, P.a, P.hslope);
    const double r = pt.r;
    const double irho2 = drcp(r * r + P.a * P.a * (pt.c * pt.c));
    const double gcon00 = -1. - 2. * r * irho2, gcon01 = 2. * irho2;   // metric.c:207-231
    const double V1 = v[2], V2 = v[3], V3 = v[4];
    const double VdotV = g.g11 * V1 * V1 + 2. * g.g13 * V1 * V3 + g.g22 * V2 * V2 + g.g33 * V3 * V3;
    const double Vfac = dsqrt(-drcp(gcon00) * (1. + fabs(VdotV)));
    const double u[4] = {-Vfac * gcon00, V1 - Vfac * gcon01, V2, V3};
    double ud[4];
    lower(g, u, ud);
    const double UdotBp = ud[1] * v[5] + ud[2] * v[6] + ud[3] * v[7];
    const double iu0 = drcp(u[0]);
    const double b[4] = {UdotBp, (v[5] + u[1] * UdotBp) * iu0, (v[6] + u[2] * UdotBp) * iu0,
                         (v[7] + u[3] * UdotBp) * iu0};
    double bd[4];
    lower(g, b, bd);
    const double Bsq = b[0] * bd[0] + b[1] * bd[1] + b[2] * bd[2] + b[3] * bd[3];
    F.ne = rho * P.Ne_unit;
    F.B = dsqrt(Bsq) * P.B_unit + 1e-20;
    const double beta = uu * (4. / 3. - 1.) * drcp(0.5 * (Bsq + 1e-20));
    const double b2 = beta * beta;
    const double w = drcp(1. + b2);
    const double trat = P.R_high * b2 * w + P.R_low * w;
    const double itr2 = drcp(trat + 2.), itr1 = drcp(1. + trat);
    const double gam2 = 0.5 * ((1. + 2. / 3. * (trat + 1.) * itr2) + 4. / 3.);
    const double mr = kProtonMass / kElectronMass;
    const double uor = uu * drcp(rho);
    F.thetae = uor * ((gam2 - 1.) * mr * itr1);
    const double dgam2 = (1. / 3.) * (itr2 * itr2);
    F.dthetae = uor * (mr * (dgam2 * (1. + trat) - (gam2 - 1.)) * (itr1 * itr1)) * (b2 * w);
    if (F.thetae < 0.) {
        F.thetae = 1e-20;
        F.dthetae = 0.;
    }
    if (Bsq * drcp(rho) > 1.)
        return false;
    double kd[4];
    lower(g, y + 4, kd);
    const double ku = kd[0] * u[0] + kd[1] * u[1] + kd[2] * u[2] + kd[3] * u[3];
    double mu = 0.;
    if (F.B != 0.) {
        mu = (kd[0] * b[0] + kd[1] * b[1] + kd[2] * b[2] + kd[3] * b[3]) *
             drcp(fabs(ku) * (F.B / P.B_unit));
        if (fabs(mu) > 1.)
            mu /= fabs(mu);
    }
    F.mu = mu;
    F.gdop = -ku;
    return true;
}

// radiative_transfer.c:71-98 for one sample in FP64 with the thermal fit
// (j_nu.c:130-153) and Kirchhoff's absorptivity with the Planck function
// (j_nu.c:188-200) as the reference writes it, exp(x) - 1: at radio
// frequencies x = h nu / k T_e is ~1e-9 and that form is off by ~1e-16 / x
// relative, which this path reproduces, since it exists to reach the gate
// decisions of the reference's FP64 optical depth.  tau lives in the
// accumulator's (tau, ctau) pair; I and, with GRAD, the tangents accumulate
// as in transfer_sample (accumulator f: the render's frequency f0 + f).
// Returns false where the sigma cut skips the sample.
// One frequency's terms of an FP64 sample: its optical depth dtau and the
// factors of its intensity dI = j exp(-tau) dl gi2 (tau after the sample),
// with the logarithmic derivatives the tangents need.
struct SampleF64 {
    double dtau, j, dl, gi2;
    double dlnj0, dlnj1, dlna1;
};

// The per-sample quantities of the thermal fit in FP64 (set once a sample).
struct SynchF64 {
    double nu_s, xk;
    bool on;
};

__host__ __device__ inline SynchF64 synch_f64(const Fluid64 &F)
{
    SynchF64 S;
    const double sth = dsqrt(1. - F.mu * F.mu);
    const double nu_c = kElectronCharge * F.B / (2. * kPi * kElectronMass * kC);
    S.nu_s = (2. / 9.) * nu_c * F.thetae * F.thetae * sth;
    S.on = !(fabs(sth) < 1.e-10) && F.thetae > 1e-4;
    S.xk = kPlanck / (kElectronMass * kC * kC) * drcp(F.thetae);   // x = h nu_p / k T_e
    return S;
}

template <bool GRAD>
__host__ __device__ inline SampleF64 sample_terms_f64(const Fluid64 &F, const SynchF64 &S,
                                                      double nu, double dl)
{
    const double c1112 = 1.8877486253633868;   // 2^(11/12)
    SampleF64 T;
    double nu_p = F.gdop * nu;
    if (nu_p <= 0.)
        nu_p = 1.;
    const double x = S.xk * nu_p;
    double j = 0., s6 = 0.;
    if (S.on && nu_p != 1.) {
        const double X = nu_p * drcp(S.nu_s);
        s6 = dexp(dlog(X) * (1. / 6.));        // X^(1/6)
        const double ff = s6 * s6 * s6 + c1112 * s6;
        j = F.ne * (1.4142135623730951 * kPi / 3.) * S.nu_s * drcp(2. * F.thetae * F.thetae) *
            (ff * ff) * dexp(-(s6 * s6)) * kE2C;
    }
    // alpha = j (e^2/ch) / B_nu / (e^2/c) with B_nu = 2 nu^3 / c^2 / (e^x - 1),
    // written as a product: where e^x - 1 is 0 (B_nu infinite) it is 0, and
    // where e^x overflows (B_nu 0) it is infinite, or NaN for j = 0 -- the
    // reference's divisions give the same
    const double alpha = j * (kE2CH / kE2C) * (dexp(x) - 1.) * (0.5 * kC * kC) *
                         drcp(nu_p * nu_p * nu_p);
    T.dtau = nu_p * drcp(nu) * alpha * dl;
    const double gi = nu * drcp(nu_p);
    T.j = j;
    T.dl = dl;
    T.gi2 = gi * gi;
    T.dlnj0 = T.dlnj1 = T.dlna1 = 0.;
    if (GRAD && j != 0.) {
        const double cb = s6 * s6, sx = cb * s6;
        const double Dx = 2. * (0.5 * sx + c1112 / 6. * s6) * drcp(sx + c1112 * s6) - cb / 3.;
        const double dlnB = 0.5 * (F.B - 1e-20) * drcp(F.B);
        const double dlnT = F.dthetae * drcp(F.thetae);
        const double dlnBnu = x * drcp(-dexpm1(-x));
        T.dlnj0 = 1. + dlnB * (1. - Dx);
        T.dlnj1 = -2. * Dx * dlnT;
        T.dlna1 = T.dlnj1 - dlnBnu * dlnT;
    }
    return T;
}

// A sample's terms into frequency f of the accumulator, its optical depth
// already at tau (after the sample; etau = exp(-tau)).
template <int NF, bool GRAD>
__host__ __device__ inline void add_sample_f64(Accum<NF, GRAD> &acc, int f, const SampleF64 &T,
                                               double tau, double etau)
{
    const double lnth = 6.907755278982137;   // log(1000.)
    const double dI = T.j * etau * T.dl * T.gi2;
    if (GRAD && T.dtau != 0.) {
        acc.dtau[f][0] += T.dtau * T.dlnj0;
        acc.dtau[f][1] += T.dtau * T.dlna1;
    }
    if (tau < lnth && !(tau < 0.)) {
        acc.I[f] += (float)dI;
        if (GRAD && dI != 0.) {
            acc.dI[f][0] += dI * (T.dlnj0 - acc.dtau[f][0]);
            acc.dI[f][1] += dI * (T.dlnj1 - acc.dtau[f][1]);
        }
    }
}

template <int NF, bool GRAD, bool DIM3>
__host__ __device__ inline bool transfer_sample_f64(const Params &P,
                                                    const float4 *__restrict__ prims,
                                                    const Point &pt, const double y[8],
                                                    double dlabs, Accum<NF, GRAD> &acc,
                                                    int f0 = 0)
{
    Fluid64 F;
    if (!fluid_sample_f64<DIM3>(P, prims, pt, y, F))
        return false;
    const double dl = dlabs * P.Rg;
    const SynchF64 S = synch_f64(F);
#pragma unroll
    for (int f = 0; f < NF; f++) {
        // unrolled (the accumulators stay in registers); frequencies past
        // P.nf are skipped
        if (f0 + f >= P.nf)
            continue;
        const SampleF64 T = sample_terms_f64<GRAD>(F, S, P.freq[f0 + f], dl);
        const double tau = tau_f64(acc.tau[f], acc.ctau[f]) + T.dtau;
        set_tau_f64(acc.tau[f], acc.ctau[f], tau);
        add_sample_f64(acc, f, T, tau, dexp(-tau));
    }
    return true;
}

// A ray whose optical depth has passed the gate at every frequency adds
// nothing more to its pixel: tau only grows (dtau >= 0, or NaN, after which
// no sample adds either) and a sample adds to I and its tangents only while
// tau < ln 1000 (radiative_transfer.c:90-95).  Such a ray ends there
// (Ray::terminate), with the outputs of a trace to the horizon or the outer
// edge.  The FP32 transfer ends 2 kGateMargin past the gate, beyond the gate
// window (its Kahan-summed tau can step back by an ulp or two); the FP64 one
// at the gate.
constexpr float kOpaqueTau = 6.907755278982137f + 2.f * kGateMargin;

// The smallest optical depth over the accumulators against the gate: one
// comparison instead of one per frequency (their AND chain held a mask
// register pair per frequency, which at 8 frequencies pushed the trace
// kernel into spilling scalar registers).  Accumulators past the render's
// frequencies repeat its last frequency (fill_params), so they change
// nothing.  fminf ignores a NaN operand, so a frequency whose tau is NaN
// (after which no sample adds, like past the gate) does not keep the ray
// going, and all-NaN gives NaN, which ends it.
template <int NF, bool GRAD>
__host__ __device__ inline bool opaque(const Params &P, const Accum<NF, GRAD> &acc)
{
    float m = acc.tau[0];
#pragma unroll
    for (int f = 1; f < NF; f++)
        m = fminf(m, acc.tau[f]);
    return !(m < kOpaqueTau);
}

template <int NF, bool GRAD>
__host__ __device__ inline bool opaque_f64(const Params &P, const Accum<NF, GRAD> &acc)
{
    bool o = true;
#pragma unroll
    for (int f = 0; f < NF; f++)
        o = o && (f >= P.nf || !(tau_f64(acc.tau[f], acc.ctau[f]) < 6.907755278982137));
    return o;
}

// A ray being traced.
template <int NF, bool GRAD>
struct Ray {
    double y[8];
    Point pt;
    int thetadot_sign;   // sign of k^theta at the previous step
    int beta_sign;       // sign of the impact parameter beta
    int steps, turns, samples;
    bool terminate;
    Accum<NF, GRAD> acc;
};

// Whether a ray traced with the FP32 transfer is to be traced again with the
// FP64 one: its FP32 tau passed the gate within kGateMargin at samples
// carrying more than kGateFlip of its intensity, summed over its frequencies
// (the thermal fit, whose FP64 form is above).
template <int EMIS, int NF, bool GRAD>
__host__ __device__ inline bool needs_exact(const Ray<NF, GRAD> &R)
{
    return EMIS == kThSynch && R.acc.wrel > kGateFlip;
}

__host__ __device__ inline void pixel_impact(const Params &P, int pix, double &alpha,
                                             double &beta)
{
    int yy = pix / P.width, x = pix - yy * P.width;
    double stepx = P.cam_x / (double)P.width;
    double stepy = P.cam_y / (double)P.height;
    alpha = -P.cam_x * 0.5 + (x + 0.5) * stepx;
    beta = -P.cam_y * 0.5 + (yy + 0.5) * stepy;
}

template <int NF, bool GRAD>
__host__ __device__ inline void ray_init(const Params &P, int cam, int pix, Ray<NF, GRAD> &R)
{
    double alpha, beta;
    pixel_impact(P, pix, alpha, beta);
    init_photon(P.cam[cam], alpha, beta, R.y);
    R.pt = make_point(R.y[1], R.y[2], P.hslope);
    R.thetadot_sign = 0;
    R.beta_sign = (beta > 0.) - (beta < 0.);
    R.steps = 0;
    R.turns = 0;
    R.samples = 0;
    R.terminate = false;
    R.acc.clear();
}

// Loop condition of integrate_geodesic (integrator.c:228-229).
template <int NF, bool GRAD>
__host__ __device__ inline bool ray_alive(const Params &P, const Ray<NF, GRAD> &R)
{
    return R.pt.r < kCutoffOuter && R.pt.r > P.cutoff_inner && R.steps < P.max_steps &&
           !R.terminate;
}

// The geodesic half of one iteration of integrate_geodesic's loop
// (integrator.c:243-272): turning-point count, step size, one step of the
// chosen integrator and the escape test.  Returns false when the ray escaped
// through the outer grid edge (the `break` at integrator.c:271-272); `dl` is
// the step taken, which the sample of the new point uses.
template <int INT, bool UG = false, int NF, bool GRAD>
__host__ __device__ inline bool ray_advance(const Params &P, Ray<NF, GRAD> &R, double &dl)
{
    // integrator.c:243-247: a sign change of k^theta is a turning point
    const int sgn = (R.y[6] > 0.) - (R.y[6] < 0.);
    if (R.thetadot_sign * sgn < 0 && R.steps > 2)
        R.turns += 1;
    R.thetadot_sign = sgn;
    if ((R.beta_sign < 0 && R.turns > kMaxOrder) || (R.beta_sign > 0 && R.turns > kMaxOrder + 1))
        R.terminate = true;
    dl = stepsize(R.y, P.stepsize);
    geodesic_step<INT, UG>(R.y, R.pt, dl, P.a, P.hslope);
    return !(R.y[1] > P.stopx1 && R.y[5] < 0.);
}

// The transfer gate of integrator.c:277: open intervals in X1 and X2.
__host__ __device__ inline bool in_grid(const Params &P, const double y[8])
{
    return y[1] < P.stopx1 && y[1] > P.startx1 && y[2] > P.startx2 && y[2] < P.stopx2;
}

// The rest of the iteration (integrator.c:277-284): the sample of the new
// point, when it is inside the grid, with the step's |dlambda|; EXACT: with
// the FP64 transfer (thermal fit) instead of the FP32 one.
template <int NF, bool GRAD, bool DIM3, int EMIS = kThSynch, bool EXACT = false>
__host__ __device__ inline void ray_sample(const Params &P, const float4 *__restrict__ prims,
                                           Ray<NF, GRAD> &R, double dl)
{
    if (in_grid(P, R.y)) {
        if constexpr (EXACT) {
            static_assert(EMIS == kThSynch, "the FP64 transfer is the thermal fit's");
            if (transfer_sample_f64<NF, GRAD, DIM3>(P, prims, R.pt, R.y, fabs(dl), R.acc)) {
                R.samples++;
                R.terminate = R.terminate || opaque_f64(P, R.acc);
            }
        } else {
            Fluid F;
            if (fluid_sample<DIM3>(P, prims, R.pt, R.y, F)) {
                R.samples++;
                transfer_sample<NF, GRAD, EMIS>(P, F, fabs(dl), R.acc);
                R.terminate = R.terminate || opaque(P, R.acc);
            }
        }
    }
    R.steps++;
}

// One iteration of integrate_geodesic's loop (integrator.c:243-284).
// Returns false when the ray escaped through the outer grid edge.
template <int NF, bool GRAD, bool DIM3, int INT = kRK2, int EMIS = kThSynch, bool EXACT = false,
          bool UG = false>
__host__ __device__ inline bool ray_step(const Params &P, const float4 *__restrict__ prims,
                                         Ray<NF, GRAD> &R)
{
    double dl;
    if (!ray_advance<INT, UG>(P, R, dl))
        return false;
    ray_sample<NF, GRAD, DIM3, EMIS, EXACT>(P, prims, R, dl);
    return true;
}

// The second trace of a ray that needs_exact, from its start state in R:
// the FP32 transfer again, now with each frequency's share of the samples in
// the gate window in wf, beside the FP64 transfer into x on the same
// geodesic.  A frequency whose own share exceeds kGateFlip -- the decision a
// launch of that frequency alone makes -- takes x's outputs (FP64 tau and
// tangents, intensity summed in FP32), the others keep the FP32 ones (R.acc,
// the same as the first trace's), so that batched and single-frequency
// launches give the same bits.  The accumulators' frequency f is the
// render's frequency f0 + f (exact_kernel traces one frequency per lane).
// WF = false: the decision is known already (the first trace's window
// share of this frequency alone, when it was the only one in the window),
// so only the FP64 transfer is traced, to its own crossing.
template <int NF, bool GRAD, bool DIM3, int INT, bool WF = true, bool OVL = false>
__host__ __device__ inline void exact_retrace(const Params &P, const float4 *__restrict__ prims,
                                              Ray<NF, GRAD> &R, Accum<NF, GRAD> &x,
                                              float (&wf)[NF], int f0 = 0)
{
    x.clear();
#pragma unroll
    for (int f = 0; f < NF; f++)
        wf[f] = 0.f;
    if constexpr (OVL) {
        // RK2 rays inside the grid (exact_kernel's resumed rays, a lane
        // each): the new point's samples are evaluated whether or not it is
        // inside (kept only if it is) beside the next step's first
        // acceleration, so that the two FP64 chains interleave in one lane;
        // the same arithmetic as below
        static_assert(INT == kRK2, "the overlap is RK2's");
        double A0[4];
        geodesic_accel(R.pt, P.a, P.hslope, R.y + 4, A0);
        while (ray_alive<NF, GRAD>(P, R)) {
            double dl;
            if (!ray_advance_from(P, R, dl, A0))
                break;
            const bool inside = in_grid(P, R.y);
            Accum<NF, GRAD> xt = x;
            transfer_sample_f64<NF, GRAD, DIM3>(P, prims, R.pt, R.y, fabs(dl), xt, f0);
            [[maybe_unused]] bool s32 = false;
            [[maybe_unused]] Accum<NF, GRAD> at = R.acc;
            [[maybe_unused]] float wt[NF];
            if constexpr (WF) {
                Fluid F;
                s32 = fluid_sample<DIM3, true>(P, prims, R.pt, R.y, F);
#pragma unroll
                for (int f = 0; f < NF; f++)
                    wt[f] = wf[f];
                transfer_sample<NF, GRAD, kThSynch>(P, F, fabs(dl), at, wt, f0);
            }
            geodesic_accel(R.pt, P.a, P.hslope, R.y + 4, A0);
            if (inside) {
                if constexpr (WF) {
                    if (s32) {
                        R.samples++;
                        R.acc = at;
#pragma unroll
                        for (int f = 0; f < NF; f++)
                            wf[f] = wt[f];
                    }
                }
                x = xt;
                R.terminate = R.terminate || ((!WF || opaque(P, R.acc)) && opaque_f64(P, x));
            }
            R.steps++;
        }
        return;
    }
    while (ray_alive<NF, GRAD>(P, R)) {
        double dl;
        if (!ray_advance<INT>(P, R, dl))
            break;
        if (in_grid(P, R.y)) {
            if constexpr (WF) {
                Fluid F;
                if (fluid_sample<DIM3>(P, prims, R.pt, R.y, F)) {
                    R.samples++;
                    transfer_sample<NF, GRAD, kThSynch>(P, F, fabs(dl), R.acc, wf, f0);
                }
            }
            transfer_sample_f64<NF, GRAD, DIM3>(P, prims, R.pt, R.y, fabs(dl), x, f0);
            R.terminate = R.terminate || ((!WF || opaque(P, R.acc)) && opaque_f64(P, x));
        }
        R.steps++;
    }
}

// The camera-to-grid leg of a ray (integrate_geodesic, integrator.c:228-277,
// up to the first step that lands inside the grid): no sample happens before
// that step.  Returns true when the ray reached the grid; `save(R)` was then
// last called with the ray as it was before that step, so that resuming from
// the saved ray repeats the step with the same arithmetic and goes on.
// Returns false when the ray ended first (escaped or out of steps; R as the
// loop left it).  The ray is saved only before steps that start at X1 <=
// `near`: with near = stopx1 + 4 STEPSIZE (a first-order step moves X1 by at
// most STEPSIZE) these are the last few steps of the leg.  Should a step from
// farther out land inside the grid, the ray is traced again from the camera
// and saved before every step.
template <int INT, class Save>
__host__ __device__ inline bool approach_ray(const Params &P, int cam, int pix, double near,
                                             Ray<1, false> &R, Save &&save)
{
    for (int pass = 0; pass < 2; pass++) {
        ray_init<1, false>(P, cam, pix, R);
        bool missed = false;
#pragma clang loop unroll(disable)
        while (ray_alive<1, false>(P, R)) {
            const bool saved = pass == 1 || R.y[1] <= near;
            if (saved)
                save(R);
            double dl;
            if (!ray_advance<INT, true>(P, R, dl))
                break;
            if (in_grid(P, R.y)) {
                if (saved)
                    return true;
                missed = true;
                break;
            }
            R.steps++;
        }
        if (!missed)
            return false;
    }
    return false;
}

// ray_advance for RK2 with the start point's acceleration given.
template <int NF, bool GRAD>
__host__ __device__ inline bool ray_advance_from(const Params &P, Ray<NF, GRAD> &R, double &dl,
                                                 const double A0[4])
{
    const int sgn = (R.y[6] > 0.) - (R.y[6] < 0.);
    if (R.thetadot_sign * sgn < 0 && R.steps > 2)
        R.turns += 1;
    R.thetadot_sign = sgn;
    if ((R.beta_sign < 0 && R.turns > kMaxOrder) || (R.beta_sign > 0 && R.turns > kMaxOrder + 1))
        R.terminate = true;
    dl = stepsize(R.y, P.stepsize);
    rk2_step_from(R.y, R.pt, dl, P.a, P.hslope, A0);
    return !(R.y[1] > P.stopx1 && R.y[5] < 0.);
}

// One RK2 iteration of integrate_geodesic's loop for rays inside the grid
// (the persistent kernel resuming saved rays, whose steps nearly all sample):
// the fluid sample of the new point is evaluated whether or not the point is
// inside the grid (its result is used only if it is) so that it shares a
// basic block with the next step's first acceleration, A0 on return, and the
// two independent chains interleave.  The same arithmetic as ray_step.
template <int NF, bool GRAD, bool DIM3, int EMIS = kThSynch>
__host__ __device__ inline bool ray_step_rk2_overlap(const Params &P,
                                                     const float4 *__restrict__ prims,
                                                     Ray<NF, GRAD> &R, double A0[4])
{
    double dl;
    if (!ray_advance_from(P, R, dl, A0))
        return false;
    const bool inside = in_grid(P, R.y);
    Fluid F;
    const bool sampled = fluid_sample<DIM3, true>(P, prims, R.pt, R.y, F);
    geodesic_accel(R.pt, P.a, P.hslope, R.y + 4, A0);
    if (inside && sampled) {
        R.samples++;
        transfer_sample<NF, GRAD, EMIS>(P, F, fabs(dl), R.acc);
        R.terminate = R.terminate || opaque(P, R.acc);
    }
    R.steps++;
    return true;
}

// ray_step and ray_step_rk2_overlap with the transfer chosen per lane: the
// FP64 one (transfer_sample_f64) where `exact`, the FP32 one elsewhere; the
// geodesic step is shared, so a wave with a lane in its second trace
// (needs_exact) runs the FP64 sample beside the others' FP32 sample.  The
// FP64 branch is marked unlikely (about 1 ray in 3000 takes it) and takes
// the next step's first acceleration after its sample, so that neither the
// FP32 fluid state nor that acceleration is live across the FP64 sample's
// registers: the register allocator then puts its spills in that branch
// (the 3D image kernel at 1024^2: 8.5 -> 7.7 ms).  Same arithmetic as
// before, the same bits.
template <int NF, bool GRAD, bool DIM3, int INT, bool UG = false>
__host__ __device__ inline bool ray_step_mixed(const Params &P, const float4 *__restrict__ prims,
                                               Ray<NF, GRAD> &R, bool exact)
{
    double dl;
    if (!ray_advance<INT, UG>(P, R, dl))
        return false;
    if (in_grid(P, R.y)) {
        if (__builtin_expect(exact, 0)) {
            if (transfer_sample_f64<NF, GRAD, DIM3>(P, prims, R.pt, R.y, fabs(dl), R.acc)) {
                R.samples++;
                R.terminate = R.terminate || opaque_f64(P, R.acc);
            }
        } else {
            Fluid F;
            if (fluid_sample<DIM3>(P, prims, R.pt, R.y, F)) {
                R.samples++;
                transfer_sample<NF, GRAD, kThSynch>(P, F, fabs(dl), R.acc);
                R.terminate = R.terminate || opaque(P, R.acc);
            }
        }
    }
    R.steps++;
    return true;
}

template <int NF, bool GRAD, bool DIM3>
__host__ __device__ inline bool ray_step_rk2_overlap_mixed(const Params &P,
                                                           const float4 *__restrict__ prims,
                                                           Ray<NF, GRAD> &R, double A0[4],
                                                           bool exact)
{
    double dl;
    if (!ray_advance_from(P, R, dl, A0))
        return false;
    const bool inside = in_grid(P, R.y);
    if (__builtin_expect(exact, 0)) {
        if (inside && transfer_sample_f64<NF, GRAD, DIM3>(P, prims, R.pt, R.y, fabs(dl), R.acc)) {
            R.samples++;
            R.terminate = R.terminate || opaque_f64(P, R.acc);
        }
        geodesic_accel(R.pt, P.a, P.hslope, R.y + 4, A0);
    } else {
        Fluid F;
        const bool sampled = fluid_sample<DIM3, true>(P, prims, R.pt, R.y, F);
        geodesic_accel(R.pt, P.a, P.hslope, R.y + 4, A0);
        if (inside && sampled) {
            R.samples++;
            transfer_sample<NF, GRAD, kThSynch>(P, F, fabs(dl), R.acc);
            R.terminate = R.terminate || opaque(P, R.acc);
        }
    }
    R.steps++;
    return true;
}

// Host set-up.
// newtonraphson.c:49-72 solved to a fixed point (the root is unique)
inline double theta_to_x2(double theta, double h)
{
    double x = theta / kPi, prev = -1.;
    for (int it = 0; it < 200 && x != prev; it++) {
        prev = x;
        double f = kPi * x + 0.5 * (1. - h) * std::sin(2. * kPi * x) - theta;
        double fp = kPi + kPi * (1. - h) * std::cos(2. * kPi * x);
        x = x - f / fp;
        if (std::fabs(x - prev) <= 1e-15 * std::fabs(x))
            break;
    }
    return std::fmin(1., std::fmax(x, 0.));
}

// metric.c:650-692 (pixel-independent part) and GRmath.c:135-184
inline Camera make_camera(double incl_deg, double a, double h)
{
    Camera c;
    double mu0 = std::cos(incl_deg / 180. * kPi);
    double thc = std::acos(mu0);
    double sinc2 = std::sin(thc) * std::sin(thc);
    double cosc2 = std::cos(thc) * std::cos(thc);
    c.sin_i = std::sqrt(1. - mu0 * mu0);
    c.mu0sq = mu0 * mu0;
    c.cot2 = cosc2 / sinc2;
    c.cosc2 = cosc2;
    double r = kRcam;
    double st = std::sin(thc), ct = std::cos(thc);
    double sigma = r * r + a * a * ct * ct;
    double delta = r * r + a * a - 2. * r;
    double A_ = (r * r + a * a) * (r * r + a * a) - delta * a * a * st * st;
    c.g11 = sigma / delta * r * r;
    c.g00 = -A_ / (sigma * delta);
    c.g03 = -2. * a * r / (sigma * delta);
    c.g33 = (delta - a * a * st * st) / (sigma * delta * st * st);
    c.g22 = 1. / sigma;
    double X1 = std::log(kRcam);
    double rc = std::exp(X1);
    double dc = rc * rc - 2. * rc + a * a;
    c.t01 = 2. * rc / dc * rc;
    c.t31 = a / dc * rc;
    double rp = 1. + std::sqrt(1. - a * a), rm = 1. - std::sqrt(1. - a * a);
    double lg = std::log((rc - rp) / (rc - rm));
    c.X1 = X1;
    c.X3 = 90. / 180. * kPi + (a / (2. * std::sqrt(1. - a * a)) * lg);
    c.X2 = theta_to_x2(thc, h);
    c.thp_inv = 1. / (kPi * (1. + (1. - h) * std::cos(2. * kPi * c.X2)));
    return c;
}


// Host-side set-up of a launch from the C-ABI structs (grmhd.c:17-33 units,
// utilities.c:17-21 cutoff).  Returns NULL or an error message.
// The kappa = 3.5 fits' constant factors (j_nu.c:35-100), FP64 on the host.
inline KappaConst kappa_const()
{
    const double k = 3.5;
    KappaConst K;
    const double Jl = 4. * kPi * std::tgamma(k - 4. / 3.) / (std::pow(3., 7. / 3.) * std::tgamma(k - 2.));
    const double Jh = std::pow(3., (k - 1.) / 2.) * (k - 2.) * (k - 1.) / 4. *
                      std::tgamma(k / 4. - 1. / 3.) * std::tgamma(k / 4. + 4. / 3.);
    K.lJlow = (float)std::log2(Jl);
    K.lJhigh = (float)std::log2(Jh);
    K.xj = (float)(3. * std::pow(k, -1.5));
    const double Al = std::pow(3., 1. / 6.) * (10. / 41.) * std::pow(2. * kPi, 2) * (k - 2.) *
                      (k - 1.) * k / (3. * k - 1.) * std::tgamma(5. / 3.);
    const double Ah = (std::pow(kPi, 2.5) / 3.) * (k - 2.) * (k - 1.) * k *
                      (2. * std::tgamma(2. + k / 2.) / (2. + k) - 1.) *
                      (std::pow(3. / k, 19. / 4.) + 3. / 5.);
    K.lAlow = (float)std::log2(Al);
    K.lAhigh = (float)std::log2(Ah);
    K.xa = (float)std::pow(-7. / 4. + 8. * k / 5., -43. / 50.);
    const double a = k - 1. / 3., b = k + 1., c = k + 2. / 3.;
    K.ha = (float)a;
    K.hb = (float)b;
    K.hc = (float)c;
    K.g1 = (float)(std::tgamma(c) * std::tgamma(b - a) / (std::tgamma(b) * std::tgamma(c - a)));
    K.g2 = (float)(std::tgamma(c) * std::tgamma(a - b) / (std::tgamma(a) * std::tgamma(c - b)));
    K.tab = nullptr;
    K.tu0 = 0.f;
    K.tidu = 1.f;
    K.tn = 0;
    return K;
}

// The absorptivity's hypergeometric factor in FP64, by the same two forms as
// kappa_hyp (Pfaff's transformation for -1 <= z < 0, the 1/(1-z)
// continuation below), each series summed to 1e-17.
inline double kappa_hyp_f64(double z)
{
    const double k = 3.5, a = k - 1. / 3., b = k + 1., c = k + 2. / 3.;
    auto series = [](double pa, double pb, double pc, double x) {
        double term = 1., sum = 1.;
        for (int n = 0; n < 2000; n++) {
            term *= (pa + n) * (pb + n) / ((pc + n) * (n + 1.)) * x;
            sum += term;
            if (std::fabs(term) <= 1e-17 * std::fabs(sum))
                break;
        }
        return sum;
    };
    if (z >= -1.)
        return std::pow(1. - z, -a) * series(a, c - b, c, z / (z - 1.));
    const double g1 = std::tgamma(c) * std::tgamma(b - a) / (std::tgamma(b) * std::tgamma(c - a));
    const double g2 = std::tgamma(c) * std::tgamma(a - b) / (std::tgamma(a) * std::tgamma(c - b));
    return std::pow(-z, -a) * g1 +
           std::pow(1. - z, -b) * g2 * series(b, c - a, b - a + 1., 1. / (1. - z));
}

// The table kappa_lhyp reads: log2 kappa_hyp(-2^u) and its derivative in u
// (central difference, step 1e-4, in FP64) at u = -32 + i/16, i < 1025.
inline void kappa_table(std::vector<float2> &tab, float &u0, float &idu)
{
    const int n = 1025;
    u0 = -32.f;
    idu = 16.f;
    tab.resize(n);
    auto g = [](double u) { return std::log2(kappa_hyp_f64(-std::exp2(u))); };
    for (int i = 0; i < n; i++) {
        const double u = -32. + i / 16.;
        tab[i].x = (float)g(u);
        tab[i].y = (float)((g(u + 1e-4) - g(u - 1e-4)) / 2e-4);
    }
}

template <class RtParams, class RtGrid>
inline const char *fill_params(const RtParams &p, const RtGrid &g, const double *incl, int ncam,
                               const double *freqs, int nf, Params &P)
{
    memset(&P, 0, sizeof P);
    if ((long long)g.n1 * g.n2 * g.n3 > kMaxCells)
        return "grids of more than 2^27 cells are not supported";
    P.n1 = g.n1;
    P.n2 = g.n2;
    P.n3 = g.n3;
    P.s1 = (unsigned)(2LL * g.n2 * g.n3);
    P.s2 = (unsigned)(2 * g.n3);
    P.startx1 = g.startx1;
    P.startx2 = g.startx2;
    P.startx3 = g.startx3;
    P.dx1 = g.dx1;
    P.dx2 = g.dx2;
    P.dx3 = g.dx3;
    P.idx1 = 1. / g.dx1;
    P.idx2 = 1. / g.dx2;
    P.idx3 = g.n3 > 1 ? 1. / g.dx3 : 0.;
    P.period3 = g.n3 * g.dx3;
    P.iperiod3 = g.n3 > 1 ? 1. / P.period3 : 0.;
    P.stopx1 = g.startx1 + g.n1 * g.dx1;
    P.stopx2 = g.startx2 + g.n2 * g.dx2;
    P.a = g.a;
    P.hslope = g.hslope;
    P.width = p.width;
    P.height = p.height;
    P.npix = p.width * p.height;
    P.cam_x = p.cam_size_x;
    P.cam_y = p.cam_size_y;
    P.stepsize = p.stepsize;
    P.max_steps = p.max_steps;
    P.cutoff_inner = (1. + std::sqrt(1. - g.a * g.a)) * (1. + kHorizonMarg);
    // grmhd.c:17-33
    double L_unit = kGGrav * p.MBH / (kC * kC);
    double RHO_unit = p.M_unit / std::pow(L_unit, 3);
    P.B_unit = kC * std::sqrt(4. * kPi * RHO_unit);
    P.Ne_unit = RHO_unit / (kProtonMass + kElectronMass);
    P.Rg = kGGrav * p.MBH / kC / kC;
    P.R_high = p.R_high;
    P.R_low = p.R_low;
    P.fNe_unit = (float)P.Ne_unit;
    P.fB_unit = (float)P.B_unit;
    P.fiB_unit = (float)(1. / P.B_unit);
    P.fR_high = (float)p.R_high;
    P.fR_low = (float)p.R_low;
    P.fa = (float)g.a;
    P.fpi_h = (float)((1. - g.hslope) * kPi);
    P.emission = p.emission;
    P.kc = kappa_const();
    P.nf = nf;
    for (int f = 0; f < kMaxFreq; f++) {
        double nu = freqs[std::min(f, nf - 1)];
        if (!(nu > 0.))
            return "frequencies must be positive";
        // above ~1e18 Hz the FP32 Planck factor overflows (h nu > 88 k T_e at
        // the Theta_e gate) and nu^-3 factors leave FP32's normal range
        if (nu > 1e17)
            return "frequencies above 1e17 Hz are outside the FP32 transfer's range";
        P.freq[f] = nu;
        P.fnu[f] = (float)nu;
        P.flnu6[f] = (float)((std::log2(nu) - 37.) / 6.);
        P.fkap[f] = (float)(0.5 * kE2CH * kC * kC / (nu * nu * nu));
    }
    P.ncam = ncam;
    for (int i = 0; i < ncam; i++) {
        if (!(incl[i] > 0. && incl[i] < 180.))
            return "inclinations must lie in (0, 180) degrees";
        P.cam[i] = make_camera(incl[i], g.a, g.hslope);
    }
    return nullptr;
}

}  // namespace rt
