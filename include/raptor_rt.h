/*
 * raptor_rt.h -- C-ABI of the MI355X per-pixel ray tracer.
 *
 * This library replaces the pixel loop of RAPTOR 1.0,
 *     void calculate_image(real **intensityfield, real energy_spectrum[],
 *                          real frequencies[])          (functions.h:70, core.c:74-152)
 * and the per-sample model hook behind it,
 *     int get_fluid_params(real X[4], real *Ne, real *Thetae, real *B, ...)
 *                                                      (functions.h:24-25, raptor_harm_model.c:190-320)
 * with one HIP implementation on one GPU.  Everything the reference passes
 * through globals (parameters.h:58-189, constants.h:32-35) is passed here
 * explicitly; there is no global state (beyond the device array of
 * rt_render_device_noalloc), and one context serves one host thread at a
 * time.
 *
 * Conventions
 *  - Every function returns RT_OK (0) or a negative RT_E* code; the message
 *    is available from rt_last_error(ctx) (rt_last_error(NULL) for errors of
 *    rt_create).  Nothing calls exit().
 *  - Plain pointers and sizes only.  Pointers named h_* are host memory,
 *    d_* device memory of the context's device.
 *  - Images are I_nu in cgs (erg s^-1 cm^-2 Hz^-1 sr^-1) after the nu^3
 *    scaling of core.c:136, laid out [camera][y*W + x][frequency] like
 *    intensityfield[y*W + x][f] (core.c:108-113); spectra are sums over pixels
 *    (core.c:146-150).  JANSKY_FACTOR (utilities.c:32-42) is applied by the
 *    writer, not here; rt_jansky_factor() returns it.
 *  - All W*H pixels are traced (the reference silently skips pixel indices
 *    >= 1e6, core.c:81-103).
 */
#ifndef RAPTOR_RT_H
#define RAPTOR_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RT_OK 0
#define RT_E_INVALID (-1)   /* bad argument or size */
#define RT_E_HIP (-2)       /* HIP runtime error */
#define RT_E_STATE (-3)     /* call out of order (e.g. render before load) */
#define RT_E_NOMEM (-4)     /* host or device allocation failed */

#define RT_MAX_FREQ 16
#define RT_MAX_CAMERAS 64

typedef struct rt_ctx rt_ctx;

/* HARM grid header (raptor_harm_model.c:90-135 reads it into the globals
 * startx, dx, a, hslope, R0; parameters.h:80-85).  n3 == 1 is the reference's
 * 2D case; n3 > 1 is the 3D extension (periodic in X3 = phi). */
typedef struct {
    int n1, n2, n3;
    double startx1, startx2, startx3;
    double dx1, dx2, dx3;
    double a, hslope, R0;
} rt_grid;

/* Run parameters: model.in records (core.c:22-34) and argv (core.c:37-42). */
typedef struct {
    double MBH;             /* black-hole mass, g              (model.in MBH) */
    double M_unit;          /* mass unit, g                    (argv[3])      */
    double R_high, R_low;   /* R-beta electron model           (argv[5], [6]) */
    int width, height;      /* pixels                          (IMG_WIDTH/HEIGHT) */
    double cam_size_x;      /* image size in M                 (CAM_SIZE_X/Y) */
    double cam_size_y;
    double stepsize;        /* adaptive step scale             (STEPSIZE)     */
    int max_steps;          /* step cap, 1e5 in parameters.h:182              */
    int integrator;         /* int_method (parameters.h:186-189): RT_INT_RK2
                               (0, the reference's default), RT_INT_RK4,
                               RT_INT_VERLET (integrator.c:13-105)            */
    int emission;           /* emission model: RT_EMIS_THSYNCH (0, the one
                               emission_coeff calls, j_nu.c:20-26),
                               RT_EMIS_THSYNCHAV (j_nu.c:157-167 times e^2/c,
                               Kirchhoff absorptivity) or RT_EMIS_KAPPA (the
                               kappa = 3.5 fits of j_nu.c:28-106, width
                               Theta_e).  Gradients: RT_EMIS_THSYNCH only.    */
} rt_params;

#define RT_INT_RK2 0
#define RT_INT_RK4 1
#define RT_INT_VERLET 2

#define RT_EMIS_THSYNCH 0
#define RT_EMIS_THSYNCHAV 1
#define RT_EMIS_KAPPA 2

/* Per-ray counters written when requested: RK2 steps taken and samples for
 * which the transfer was evaluated (inside the grid and past the sigma cut).
 * A ray ends once its optical depth has passed ln 1000 at every frequency
 * (no later sample reaches its pixel, radiative_transfer.c:90-95), so such a
 * ray counts fewer steps than the reference's loop takes. */
typedef struct {
    int32_t steps;
    int32_t samples;
} rt_ray_counters;

/* ---- lifetime ---------------------------------------------------------- */

/* Create a context on HIP device `device`; NULL on failure. */
rt_ctx *rt_create(int device);
void rt_destroy(rt_ctx *ctx);
const char *rt_last_error(const rt_ctx *ctx);
/* Library version string. */
const char *rt_version(void);

/* ---- set-up (once per run / dump) --------------------------------------- */

/* Replaces the globals set by read_model (core.c:4-72) and set_units
 * (grmhd.c:17-33).  May be called again between renders. */
int rt_set_params(rt_ctx *ctx, const rt_params *p);

/* Replaces init_model()'s storage step (raptor_harm_model.c:59-66,
 * grmhd.c:76-91): copies the 8 primitives KRHO, UU, U1, U2, U3, B1, B2, B3
 * (raptor_harm_model.h:67-74) to HBM once.  h_prims is float32 in the
 * reference's index order p[v][i][j][k], i.e. h_prims[((v*n1+i)*n2+j)*n3+k].
 * Requires n1, n2 >= 2, n3 >= 1, R0 == 0 (the reference's inverse metric
 * assumes it, metric.c:207-231), and at most 2^27 cells (4 GiB of FP32
 * primitives; the samplers address them with 32-bit offsets). */
int rt_load_prims(rt_ctx *ctx, const rt_grid *grid, const float *h_prims);

/* ---- the pixel loop ----------------------------------------------------- */

/* calculate_image for n_cam inclinations (degrees) and n_freq frequencies
 * (Hz, in (0, 1e17]: the FP32 transfer's range) in one launch.  h_image: [n_cam][W*H][n_freq] doubles; h_spectrum
 * (may be NULL): [n_cam][n_freq].  Synchronous. */
int rt_render(rt_ctx *ctx, const double *inclinations, int n_cam,
              const double *freqs, int n_freq, double *h_image, double *h_spectrum);

/* Same, plus the image's derivatives with respect to ln(M_unit) and R_high:
 * h_dimage [n_cam][W*H][n_freq][2] (index 0: d/d ln M_unit, 1: d/d R_high). */
int rt_render_grad(rt_ctx *ctx, const double *inclinations, int n_cam,
                   const double *freqs, int n_freq, double *h_image, double *h_dimage,
                   double *h_spectrum);

/* Device-resident variant: writes d_image (and d_dimage if non-NULL,
 * d_counters [n_cam*W*H] if non-NULL, and d_spectrum [n_cam][n_freq] -- the
 * energy spectrum of core.c:146-150, the sum of each image over its pixels --
 * if non-NULL) in HBM on the context's stream and returns without
 * synchronising.  The buffers must hold the sizes above. */
int rt_render_device(rt_ctx *ctx, const double *inclinations, int n_cam,
                     const double *freqs, int n_freq, double *d_image, double *d_dimage,
                     rt_ray_counters *d_counters, double *d_spectrum);

/* rt_render_device without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and outputs, bit for bit.  The work queue,
 * the FP64 re-trace's ray list and the spectrum's partial sums live in one
 * static device array of the library (34 MiB per GPU, held until the
 * process ends) instead of buffers the library allocates; hence at most
 * RT_NOALLOC_MAX_RAYS work items per call (n_cam x W x H, in whole claim
 * tiles), RT_E_INVALID beyond.  It traces every ray from the camera in one
 * launch (the launch configuration's approach is not used) and claims rays
 * in the centre-out order only if an earlier rt_render_device on this
 * context built that order for the image shape, in tiled scan order
 * otherwise.  Every call on one GPU uses that GPU's array: calls that
 * overlap in time on one GPU (from several contexts or threads) race and
 * give wrong results. */
#define RT_NOALLOC_MAX_RAYS (1 << 22)
int rt_render_device_noalloc(rt_ctx *ctx, const double *inclinations, int n_cam,
                             const double *freqs, int n_freq, double *d_image,
                             double *d_dimage, rt_ray_counters *d_counters,
                             double *d_spectrum);

/* Wait for the context's stream. */
int rt_synchronize(rt_ctx *ctx);

/* Duration in ms of the last render's trace kernel(s), from HIP events on
 * the context's stream (valid after rt_synchronize). */
double rt_last_kernel_ms(const rt_ctx *ctx);

/* Durations in ms of the last eager render's kernels, from HIP events on the
 * context's stream (synchronises it): out[0] the approach kernel with the
 * queue reset (about 0 when no approach ran), out[1] the persistent trace
 * kernel, out[2] the FP64 re-trace of the rays at the tau gate, out[3] the
 * spectrum reduction (about 0 when none was asked for), out[4] the ring
 * order's lead kernel, which runs beside the others on a second stream (0
 * when none ran).  out holds 5 doubles.  Kept for renders of one launch
 * chunk (at most 2^23 rays); RT_E_STATE otherwise.  No counterpart in the
 * reference, which times its pixel loop with clock() (core.c:91-93,116). */
int rt_last_kernel_times(rt_ctx *ctx, double *out);

/* JANSKY_FACTOR of utilities.c:32-42 for the current parameters. */
double rt_jansky_factor(const rt_ctx *ctx);

/* ---- launch configuration ------------------------------------------------ */

/* Launch shape of the persistent trace kernel.  block: threads per
 * workgroup (64; other values are rejected).  blocks_per_cu: resident
 * workgroups per CU (0 = library default: the occupancy limit, or one wave
 * per SIMD for launches of fewer than 2^19 rays, which end on their longest
 * rays).  refill: a lane whose ray has
 * ended claims a new one only when at least `refill` lanes of its wave are
 * idle (1..64; 0 = library default).  tile: rays are handed out in T x T
 * pixel tiles (1 = single pixels; 0 = library default, 8).  order: the order
 * of the tiles, RT_ORDER_CENTRE (by distance from the image centre, so the
 * long rays near the photon ring start first and the short escaping rays
 * fill the end of the launch), RT_ORDER_SCAN (row-major) or RT_ORDER_RING:
 * the centre order, and for renders of one camera and one frequency that
 * run the split launch below in one chunk, the longest rays -- the pixels
 * within 0.3 M of the camera's critical curve (the photon ring) and those
 * whose rays turn in theta within arcsin 0.02 of the spin axis (where the
 * step-size rule shrinks the steps), at most 5% of the image -- are traced
 * from the camera by a lead kernel on a second stream of the context from
 * the start of the render, beside the approach kernel, instead of after it;
 * 0 = library default (ring).  approach: the share of the rays, in claim order,
 * whose camera-to-grid leg runs first in a kernel of its own: 1..100
 * percent, RT_APPROACH_NONE for none, 0 = library default (all rays for
 * launches of at least 2^20 rays, none below).  Those rays are
 * saved where they reach the grid (128 bytes each, at most 2^23 saved rays
 * per launch pair, larger batches in chunks) and resumed first by the
 * persistent kernel, whose waves then do no camera-to-grid steps beside
 * rays inside the grid; it traces the other rays from the camera, and
 * these fill the time in which the longest resumed rays finish.  None of
 * these changes any output bit.  The centre and ring orders for a new image
 * shape (and, ring, camera), and the buffer of saved rays, are made by an
 * eager render, so render once before capturing a graph. */
#define RT_ORDER_CENTRE 1
#define RT_ORDER_SCAN 2
#define RT_ORDER_RING 3

/* The claim order that a render of one camera at `inclination` (degrees)
 * and one frequency uses with the current image parameters and launch
 * configuration: h_order[W*H] pixel indices (y*W + x) in the order their
 * rays are claimed, of which the first *n_lead are traced by the ring
 * order's lead kernel (0 when it does not run).  Builds the table if
 * needed, as an eager render would.  No counterpart in the reference, whose pixel loop runs in index
 * order (core.c:107-113): for measurement and tests. */
int rt_claim_order(rt_ctx *ctx, double inclination, int *h_order, int *n_lead);
#define RT_APPROACH_NONE (-1)
typedef struct {
    int block;
    int blocks_per_cu;
    int refill;
    int tile;
    int order;
    int approach;
} rt_launch_config;

int rt_set_launch_config(rt_ctx *ctx, const rt_launch_config *cfg);
int rt_get_launch_config(const rt_ctx *ctx, rt_launch_config *cfg);

/* ---- graphs of repeated launches ------------------------------------------ */

/* Record every rt_render_device call made between rt_capture_begin and
 * rt_capture_end into one HIP graph (no other rt_* call that allocates,
 * copies or synchronises is allowed in between).  rt_graph_launch replays it
 * on the context's stream without synchronising; rt_last_kernel_ms then
 * times the whole replay.  A graph renders the primitives that were loaded
 * when it was captured: rt_load_prims keeps that buffer alive, beside the
 * new one, until the last graph using it is destroyed.  Destroy a context's
 * graphs before the context.  No counterpart in the reference, whose pixel
 * loop is one blocking call (core.c:97-114). */
typedef struct rt_graph rt_graph;
int rt_capture_begin(rt_ctx *ctx);
int rt_capture_end(rt_ctx *ctx, rt_graph **graph);
int rt_graph_launch(rt_ctx *ctx, rt_graph *graph);
int rt_graph_destroy(rt_ctx *ctx, rt_graph *graph);

/* ---- device buffers for callers that keep outputs in HBM ---------------- */
int rt_device_alloc(rt_ctx *ctx, size_t bytes, void **d_ptr);
int rt_device_free(rt_ctx *ctx, void *d_ptr);
int rt_memcpy_d2h(rt_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);

/* ---- the model hook on the device --------------------------------------- */

/* get_fluid_params (raptor_harm_model.c:190-320) evaluated by the device
 * sampler at n points X[n][4] (MKS).  out[n][14]: Ne (cm^-3), Thetae, B (G),
 * Ucon[4], Bcon[4], the hook's return value (1, or 0 for b^2/rho > 1; then Ne,
 * Thetae and B are 1e-20 as in the reference), beta and Bern. */
int rt_fluid_params(rt_ctx *ctx, const double *h_X, int n, double *h_out);

/* Device evaluation of the path's building blocks at n points, for checking:
 *   RT_PROBE_METRIC     in X[4]          -> out g_dd[16], g_uu[16]   (metric.c:94-138, 207-231)
 *   RT_PROBE_ACCEL      in X[4], k[4]    -> out A[4] = -Gamma^m_ab k^a k^b (integrator.c:126-146)
 *   RT_PROBE_RK2        in y[8]          -> out dl, y'[8] (integrator.c:44-70, 109-122)
 *   RT_PROBE_CAMERA     in alpha, beta, inclination -> out y[8]     (metric.c:650-727)
 *   RT_PROBE_EMISSION   in B, cos(pitch), Thetae, nu_p, Ne -> out j, alpha_nu (j_nu.c:130-200)
 *   RT_PROBE_EMISSION_MODEL  in B, cos(pitch), Thetae, nu_p, Ne, model (RT_EMIS_*)
 *                            -> out j, alpha_nu of that model (j_nu.c:28-200)
 */
#define RT_PROBE_METRIC 1
#define RT_PROBE_ACCEL 2
#define RT_PROBE_RK2 3
#define RT_PROBE_CAMERA 4
#define RT_PROBE_EMISSION 5
#define RT_PROBE_EMISSION_MODEL 6
int rt_probe(rt_ctx *ctx, int kind, const double *h_in, int n, double *h_out);

#ifdef __cplusplus
}
#endif
#endif /* RAPTOR_RT_H */
