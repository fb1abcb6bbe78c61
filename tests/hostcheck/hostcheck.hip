// hostcheck.hip -- test-only CPU build of the tracer's per-ray physics.
//
// The __host__ __device__ functions of raptor_amd/csrc/rt_physics.hpp are
// compiled for the host here and driven ray by ray on the CPU, so that the
// not-gpu test suite can compare the kernel's arithmetic with the oracle
// before any GPU run.  This library is test infrastructure: the product
// (libraptor_rt.so) has no CPU path.
#include <hip/hip_runtime.h>
#include <math.h>

// Host overload of the device library's sincospi for this CPU build (the
// kernel uses the device version): sin(pi x), cos(pi x) after reducing x to
// [-1, 1], accurate to a few ulp, which is all the comparison needs.
static inline __host__ void sincospi(double x, double *s, double *c)
{
    double r = x - 2. * nearbyint(0.5 * x);
    *s = sin(M_PI * r);
    *c = cos(M_PI * r);
}

#include "../../include/raptor_rt.h"
#include "../../raptor_amd/csrc/rt_physics.hpp"

#include <type_traits>
#include <vector>

using namespace rt;

// Traces R from its start state to the end; a ray that needs_exact is started
// again by `restart` and traced with the FP64 transfer, as the kernel does.
template <int NF, bool GRAD, int INT, int EMIS, bool EXACT, class Restart>
static void trace_to_end(const Params &P, const float4 *prims, Ray<NF, GRAD> &R,
                         Restart &&restart, bool retrace = true)
{
    const bool dim3 = P.n3 > 1;
    while (ray_alive<NF, GRAD>(P, R)) {
        bool cont = dim3 ? ray_step<NF, GRAD, true, INT, EMIS, EXACT>(P, prims, R)
                         : ray_step<NF, GRAD, false, INT, EMIS, EXACT>(P, prims, R);
        if (!cont)
            break;
    }
    if constexpr (!EXACT && EMIS == kThSynch) {
        if (retrace && needs_exact<EMIS>(R)) {
            // exact_kernel: the second trace, FP64 outputs where a frequency's
            // own window share asks for them; the counters stay the first
            // trace's with several frequencies (exact_kernel writes the image
            // only) and are the second's with one (the trace kernel's lane
            // traces it again itself)
            if (P.nf == 1) {
                restart(R);
                Accum<NF, GRAD> x;
                float wf[NF];
                if (P.n3 > 1)
                    exact_retrace<NF, GRAD, true, INT>(P, prims, R, x, wf);
                else
                    exact_retrace<NF, GRAD, false, INT>(P, prims, R, x, wf);
                if (wf[0] > kGateFlip) {
                    R.acc.I[0] = x.I[0];
                    if (GRAD) {
                        R.acc.dI[0][0] = x.dI[0][0];
                        R.acc.dI[0][1] = x.dI[0][1];
                    }
                }
                return;
            }
            // several frequencies: one second trace per frequency that had a
            // sample in the gate window, with that frequency alone, as
            // exact_kernel's lanes do
            const unsigned mask = R.acc.gmask;
            for (int f = 0; f < P.nf; f++) {
                if (!((mask >> f) & 1u))
                    continue;
                Ray<NF, GRAD> S;
                restart(S);
                Ray<1, GRAD> R1;
                for (int q = 0; q < 8; q++)
                    R1.y[q] = S.y[q];
                R1.pt = S.pt;
                R1.thetadot_sign = S.thetadot_sign;
                R1.beta_sign = S.beta_sign;
                R1.steps = S.steps;
                R1.turns = S.turns;
                R1.samples = 0;
                R1.terminate = false;
                R1.acc.clear();
                Accum<1, GRAD> x;
                float wf[1];
                const bool alone = (mask & (mask - 1)) == 0;
                if (alone && P.n3 > 1)
                    exact_retrace<1, GRAD, true, INT, false>(P, prims, R1, x, wf, f);
                else if (alone)
                    exact_retrace<1, GRAD, false, INT, false>(P, prims, R1, x, wf, f);
                else if (P.n3 > 1)
                    exact_retrace<1, GRAD, true, INT>(P, prims, R1, x, wf, f);
                else
                    exact_retrace<1, GRAD, false, INT>(P, prims, R1, x, wf, f);
                if (alone || wf[0] > kGateFlip) {
                    R.acc.I[f] = x.I[0];
                    if (GRAD) {
                        R.acc.dI[f][0] = x.dI[0][0];
                        R.acc.dI[f][1] = x.dI[0][1];
                    }
                }
            }
        }
    }
}

// Transfer of the thermal fit, for tests (hc_render_mode): 0 as the kernels
// (FP32, rays at the tau gate again in FP64), 1 every ray with the FP64
// transfer, 2 FP32 only (no second trace).
static int g_mode = 0;

template <int NF, bool GRAD, int INT, int EMIS, class Restart>
static void finish(const Params &P, const float4 *prims, int ray, Ray<NF, GRAD> &R, double *img,
                   double *dimg, int *cnt, Restart &&restart)
{
    if constexpr (EMIS == kThSynch) {
        if (g_mode == 1)
            trace_to_end<NF, GRAD, INT, EMIS, true>(P, prims, R, restart);
        else if (g_mode == 2)
            trace_to_end<NF, GRAD, INT, EMIS, false>(P, prims, R, [](Ray<NF, GRAD> &) {},
                                                     false);
        else
            trace_to_end<NF, GRAD, INT, EMIS, false>(P, prims, R, restart);
    } else {
        trace_to_end<NF, GRAD, INT, EMIS, false>(P, prims, R, restart);
    }
    for (int f = 0; f < P.nf; f++) {
        img[(size_t)ray * P.nf + f] = R.acc.I[f];
        if (GRAD) {
            dimg[((size_t)ray * P.nf + f) * 2] = R.acc.dI[f][0];
            dimg[((size_t)ray * P.nf + f) * 2 + 1] = R.acc.dI[f][1];
        }
    }
    if (cnt) {
        cnt[2 * ray] = R.steps;
        cnt[2 * ray + 1] = R.samples;
    }
}

// split == 0: every ray traced in one loop from the camera (the fused
// launch); otherwise the split launch's path: approach_ray with near =
// stopx1 + near_off, the saved ray copied field by field as the trace kernel
// loads it, then the same loop.
template <int NF, bool GRAD, int INT, int EMIS>
static void run(const Params &P, const float4 *prims, int total, int split, double near_off,
                double *img, double *dimg, int *cnt)
{
    for (int ray = 0; ray < total; ray++) {
        Ray<NF, GRAD> R;
        if (!split) {
            auto init = [&](Ray<NF, GRAD> &Q) { ray_init<NF, GRAD>(P, ray / P.npix, ray % P.npix, Q); };
            init(R);
            finish<NF, GRAD, INT, EMIS>(P, prims, ray, R, img, dimg, cnt, init);
            continue;
        }
        Ray<1, false> A, S;
        bool reached = approach_ray<INT>(P, ray / P.npix, ray % P.npix, P.stopx1 + near_off, A,
                                         [&](const Ray<1, false> &r) { S = r; });
        if (!reached) {
            for (int f = 0; f < P.nf; f++) {
                img[(size_t)ray * P.nf + f] = 0.;
                if (GRAD)
                    dimg[((size_t)ray * P.nf + f) * 2] = dimg[((size_t)ray * P.nf + f) * 2 + 1] = 0.;
            }
            if (cnt) {
                cnt[2 * ray] = A.steps;
                cnt[2 * ray + 1] = 0;
            }
            continue;
        }
        auto load = [&](Ray<NF, GRAD> &Q) {
            for (int q = 0; q < 8; q++)
                Q.y[q] = S.y[q];
            Q.pt = S.pt;
            Q.steps = S.steps;
            Q.turns = S.turns;
            Q.thetadot_sign = S.thetadot_sign;
            Q.beta_sign = S.beta_sign;
            Q.samples = 0;
            Q.terminate = false;
            Q.acc.clear();
        };
        load(R);
        finish<NF, GRAD, INT, EMIS>(P, prims, ray, R, img, dimg, cnt, load);
    }
}

extern "C" int hc_render_split(const rt_params *p, const rt_grid *g, const float *soa,
                               const double *incl, int ncam, const double *freqs, int nf,
                               int split, double near_off, double *img, double *dimg, int *cnt);

// The kappa fit's constants with the hypergeometric table, as the library
// hands them to its kernels.
static KappaConst host_kappa()
{
    static std::vector<float2> tab;
    static float u0, idu;
    if (tab.empty())
        kappa_table(tab, u0, idu);
    KappaConst K = kappa_const();
    K.tab = tab.data();
    K.tu0 = u0;
    K.tidu = idu;
    K.tn = (int)tab.size();
    return K;
}

// log2 of the kappa absorptivity's hypergeometric factor at z = -2^u, as the
// kernels evaluate it (table inside its range, series outside)
extern "C" double hc_kappa_lhyp(double u)
{
    return (double)kappa_lhyp(host_kappa(), (float)u);
}

extern "C" int hc_render(const rt_params *p, const rt_grid *g, const float *soa,
                         const double *incl, int ncam, const double *freqs, int nf, double *img,
                         double *dimg, int *cnt)
{
    return hc_render_split(p, g, soa, incl, ncam, freqs, nf, 0, 0., img, dimg, cnt);
}

extern "C" int hc_render_split(const rt_params *p, const rt_grid *g, const float *soa,
                               const double *incl, int ncam, const double *freqs, int nf,
                               int split, double near_off, double *img, double *dimg, int *cnt)
{
    Params *P = new Params;
    const char *msg = fill_params(*p, *g, incl, ncam, freqs, nf, *P);
    if (msg) {
        delete P;
        return -1;
    }
    P->kc = host_kappa();
    size_t ncell = (size_t)g->n1 * g->n2 * g->n3;
    std::vector<float4> aos(2 * ncell);
    for (size_t c = 0; c < ncell; c++) {
        aos[2 * c] = make_float4(soa[c], soa[ncell + c], soa[2 * ncell + c], soa[3 * ncell + c]);
        aos[2 * c + 1] = make_float4(soa[4 * ncell + c], soa[5 * ncell + c], soa[6 * ncell + c],
                                     soa[7 * ncell + c]);
    }
    int total = ncam * P->npix;
    auto go = [&](auto integ) {
        constexpr int INT = decltype(integ)::value;
        if (p->emission != RT_EMIS_THSYNCH) {   // image only, like the library
            if (dimg)
                return;
            if (p->emission == RT_EMIS_KAPPA) {
                if (nf == 1) run<1, false, INT, kKappa>(*P, aos.data(), total, split, near_off, img, dimg, cnt);
                else run<16, false, INT, kKappa>(*P, aos.data(), total, split, near_off, img, dimg, cnt);
            } else {
                if (nf == 1) run<1, false, INT, kThSynchAv>(*P, aos.data(), total, split, near_off, img, dimg, cnt);
                else run<16, false, INT, kThSynchAv>(*P, aos.data(), total, split, near_off, img, dimg, cnt);
            }
            return;
        }
        if (nf == 1) {
            if (dimg) run<1, true, INT, kThSynch>(*P, aos.data(), total, split, near_off, img, dimg, cnt);
            else run<1, false, INT, kThSynch>(*P, aos.data(), total, split, near_off, img, dimg, cnt);
        } else {
            if (dimg) run<16, true, INT, kThSynch>(*P, aos.data(), total, split, near_off, img, dimg, cnt);
            else run<16, false, INT, kThSynch>(*P, aos.data(), total, split, near_off, img, dimg, cnt);
        }
    };
    if (p->integrator == RT_INT_RK4)
        go(std::integral_constant<int, kRK4>());
    else if (p->integrator == RT_INT_VERLET)
        go(std::integral_constant<int, kVerlet>());
    else
        go(std::integral_constant<int, kRK2>());
    delete P;
    return 0;
}

// The image with the thermal fit's transfer in the given mode (g_mode).
extern "C" int hc_render_mode(const rt_params *p, const rt_grid *g, const float *soa,
                              const double *incl, int ncam, const double *freqs, int nf,
                              int mode, double *img, double *dimg, int *cnt)
{
    if (p->emission != RT_EMIS_THSYNCH || mode < 0 || mode > 2)
        return -1;
    g_mode = mode;
    const int rc = hc_render_split(p, g, soa, incl, ncam, freqs, nf, 0, 0., img, dimg, cnt);
    g_mode = 0;
    return rc;
}

// The midpoint trigonometry from the start point (shifted_point) next to a
// full evaluation at the midpoint: out[0..4] shifted, out[5..9] full.
extern "C" void hc_shifted_point(double X1, double X2, double e1, double e2, double h,
                                 double *out)
{
    Point p0 = make_point(X1, X2, h);
    Point a = shifted_point(p0, X1 + e1, X2 + e2, e1, e2, h);
    Point b = make_point(X1 + e1, X2 + e2, h);
    const double va[5] = {a.r, a.s, a.c, a.s2, a.c2}, vb[5] = {b.r, b.s, b.c, b.s2, b.c2};
    for (int i = 0; i < 5; i++) {
        out[i] = va[i];
        out[5 + i] = vb[i];
    }
}

// The FP32 emission models of the transfer pass, host build: out = j_nu
// (cgs), alpha_nu.
extern "C" void hc_emission(int model, double B, double mu, double Te, double nu, double ne,
                            double *out)
{
    if (model == kThSynch) {
        EmisF e = emission_f32((float)B, (float)mu, (float)Te, (float)nu, (float)ne);
        double jx = (double)e.jh * exp(-(double)e.cb);
        out[0] = kE2C * jx;
        out[1] = e.jh != 0.f ? jx * (double)e.kap : (e.poison ? NAN : 0.);
        return;
    }
    const KappaConst K = host_kappa();
    EmisX e = model == kKappa
                  ? emission_kappa(K, (float)B, (float)mu, (float)Te, (float)nu, (float)ne)
                  : emission_thsynchav((float)B, (float)Te, (float)nu, (float)ne);
    out[0] = kE2C * (double)e.jx;
    out[1] = (model == kKappa || e.jx != 0.f) ? (double)e.alpha : (e.poison ? NAN : 0.);
}
