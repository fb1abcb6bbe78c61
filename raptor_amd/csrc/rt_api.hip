// rt_api.hip -- trace kernels and the C-ABI of include/raptor_rt.h.
//
// The trace kernel is persistent: the grid is sized to the resident
// capacity of the GPU, and every lane claims rays from one global atomic
// counter, tracing each ray to completion and claiming the next one as soon
// as its current ray ends.  Ray lengths differ by orders of magnitude
// (escaping rays ~400 RK2 steps, rays skimming the photon ring thousands),
// so a lane-granular queue keeps the CUs busy until the last ray is claimed.
//
// A ray id is cam * W*H + (y*W + x); its outputs go to image[id*nf + f].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/raptor_rt.h"
#include "rt_physics.hpp"

using namespace rt;

static constexpr int kDefaultRefill = 4;
static constexpr int kDefaultTile = 8;   // one wave's 64 lanes take one 8 x 8 tile
// Approach kernel (percent of rays) by default: all rays of launches of at
// least kApproachRays, whose persistent kernel then runs long beside its
// longest ray; none below, where the kernel lasts about as long as its
// longest ray anyway and the approach kernel would only add its own time.
static constexpr long long kApproachRays = 1LL << 20;

namespace {

// Rays that have reached the grid, saved by the approach kernel for the trace
// kernel to resume, each as it was before the step that took it into the
// grid: one 128-byte record per ray, one cache line, written and read whole
// as eight 16-byte vectors, so that a lane's refill reads its own lines
// only (with the fields of 16 neighbouring rays in one line, refills of a
// few lanes on different CUs fetched each line several times).
// ray < 0: nothing left to trace.  flags: theta-dot sign + 1 | (beta sign +
// 1) << 2.
struct alignas(16) SavedRay {
    double y[8];    // X^mu, k^mu
    double pt[5];   // r, sin theta, cos theta, sin 2 pi X2, cos 2 pi X2
    int steps, turns, ray, flags;
    int pad[2];
};
static_assert(sizeof(SavedRay) == 128, "one cache line per saved ray");
struct RayState {
    char *d;
    int n;   // slots (a multiple of 64); 0: no saved state
};
static constexpr size_t kStateBytes = sizeof(SavedRay);

__device__ inline SavedRay *state_rec(const RayState &S, int i)
{
    return (SavedRay *)S.d + i;
}

struct TraceArgs {
    const float4 *prims;
    unsigned int *queue;
    int total;          // work items of this launch
    int w0;             // index of its first work item
    int nsaved;         // its first nsaved items were saved by the approach kernel
    int refill;
    int tile, ntx;      // work-item order: T x T pixel tiles, ntx tiles per row
    int per_cam;        // work items per camera
    int unit;           // cameras are interleaved in units of this many work items
    const int *order;   // work item -> pixel within a camera, or NULL (scan)
    double *image;
    double *dimage;
    rt_ray_counters *counters;
    RayState st;        // the saved rays
    // rays to trace again with the FP64 transfer (exact_kernel), xcount of
    // them: per ray its work item's index within the launch and the mask of
    // frequencies that had a sample in the gate window (Accum::gmask)
    int2 *xlist;
    unsigned int *xcount;
};

template <int NF, bool GRAD, bool DIM3>
__device__ inline void write_ray(const Params &P, const TraceArgs &A, int ray,
                                 const Ray<NF, GRAD> &R)
{
    const size_t o = (size_t)ray * P.nf;
    for (int f = 0; f < NF; f++) {
        if (f < P.nf) {
            A.image[o + f] = (double)R.acc.I[f];
            if (GRAD) {
                A.dimage[(o + f) * 2 + 0] = R.acc.dI[f][0];
                A.dimage[(o + f) * 2 + 1] = R.acc.dI[f][1];
            }
        }
    }
    if (A.counters) {
        rt_ray_counters c;
        c.steps = R.steps;
        c.samples = R.samples;
        A.counters[ray] = c;
    }
}

// Work item -> (camera, pixel).  The cameras' work items are interleaved in
// units of A.unit (one claim tile): unit u of the launch is unit u / n_cam of
// camera u % n_cam, so every camera's first tiles -- the long rays near the
// centre -- are claimed at the start of the launch (camera after camera, the
// last camera's long rays started when most of the launch was done, and
// ended it late); a camera's trailing partial unit comes after all whole
// units.  Within a camera, r indexes its work items: with an order table
// (per_cam = W*H) the table gives the pixel.  Otherwise tile <= 1:
// row-major pixels; or T x T tiles in row-major tile order, pixels
// row-major inside a tile, and items that fall outside a ragged last tile
// row/column are skipped.
__device__ inline bool work_to_ray(const Params &P, const TraceArgs &A, int w, int &ray)
{
    const int U = A.unit, nc = P.ncam;
    const int full = A.per_cam / U;
    const int head = full * U * nc;
    int cam, r;
    if (w < head) {
        const int u = w / U;
        cam = u % nc;
        r = (u / nc) * U + (w - u * U);
    } else {
        const int rem = A.per_cam - full * U;   // > 0 here
        const int x = w - head;
        cam = x / rem;
        r = full * U + (x - cam * rem);
    }
    int pix;
    if (A.order) {
        pix = A.order[r];
    } else if (A.tile <= 1) {
        pix = r;
    } else {
        const int tt = A.tile * A.tile;
        const int t = r / tt, l = r - t * tt;
        const int x = (t % A.ntx) * A.tile + l % A.tile;
        const int y = (t / A.ntx) * A.tile + l / A.tile;
        if (x >= P.width || y >= P.height)
            return false;
        pix = y * P.width + x;
    }
    ray = cam * P.npix + pix;
    return true;
}

// Rays that end before reaching the grid (escaped or out of steps) carry no
// emission: zero outputs.
__device__ inline void write_empty(const Params &P, const TraceArgs &A, int ray, int steps)
{
    const size_t o = (size_t)ray * P.nf;
    for (int f = 0; f < P.nf; f++)
        A.image[o + f] = 0.;
    if (A.dimage)
        for (int f = 0; f < 2 * P.nf; f++)
            A.dimage[2 * o + f] = 0.;
    if (A.counters) {
        rt_ray_counters c;
        c.steps = steps;
        c.samples = 0;
        A.counters[ray] = c;
    }
}

template <int NF, bool GRAD>
__device__ inline void save_state(const RayState &S, int i, int ray, const Ray<NF, GRAD> &R)
{
    SavedRay v;
#pragma unroll
    for (int q = 0; q < 8; q++)
        v.y[q] = R.y[q];
    v.pt[0] = R.pt.r;
    v.pt[1] = R.pt.s;
    v.pt[2] = R.pt.c;
    v.pt[3] = R.pt.s2;
    v.pt[4] = R.pt.c2;
    v.steps = R.steps;
    v.turns = R.turns;
    v.ray = ray;
    v.flags = (R.thetadot_sign + 1) | ((R.beta_sign + 1) << 2);
    v.pad[0] = v.pad[1] = 0;
    *state_rec(S, i) = v;
}

__device__ inline void drop_state(const RayState &S, int i) { state_rec(S, i)->ray = -1; }

__device__ inline int state_ray(const RayState &S, int i) { return state_rec(S, i)->ray; }

template <int NF, bool GRAD>
__device__ inline void load_state(const RayState &S, int i, Ray<NF, GRAD> &R)
{
    const SavedRay v = *state_rec(S, i);
#pragma unroll
    for (int q = 0; q < 8; q++)
        R.y[q] = v.y[q];
    R.pt.r = v.pt[0];
    R.pt.s = v.pt[1];
    R.pt.c = v.pt[2];
    R.pt.s2 = v.pt[3];
    R.pt.c2 = v.pt[4];
    // the fields in ray_init's order: the two set-ups are alternatives of one
    // branch, and the compiler pairs their stores in order when it merges
    // them, so a different order would pair different fields and leave the
    // counters in scratch memory behind a selected address
    R.thetadot_sign = (v.flags & 3) - 1;
    R.beta_sign = ((v.flags >> 2) & 3) - 1;
    R.steps = v.steps;
    R.turns = v.turns;
    R.samples = 0;
    R.terminate = false;
    R.acc.clear();
}

// The approach: one thread per work item traces its ray from the camera to
// the grid (approach_ray) and leaves it saved for the trace kernel.  Every
// ray's approach is about ln(r_cam / r_out) / STEPSIZE steps with no fluid
// work, so these waves run without the idle lanes they would have beside rays
// inside the grid.  3 waves per SIMD; RK4's four accelerations per step
// would spill 40 VGPRs there (six reloads a step), and at 2 waves (205
// VGPRs) it runs 2% faster.
template <int INT>
__global__ __launch_bounds__(64, INT == kRK4 ? 2 : 3) void approach_kernel(const Params P,
                                                                          const TraceArgs A)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= A.nsaved)
        return;
    int ray;
    if (!work_to_ray(P, A, A.w0 + i, ray)) {
        drop_state(A.st, i);
        return;
    }
    Ray<1, false> R;
    auto save = [&](const Ray<1, false> &S) { save_state(A.st, i, ray, S); };
    if (approach_ray<INT>(P, ray / P.npix, ray % P.npix, P.stopx1 + 4. * P.stepsize, R, save))
        return;
    write_empty(P, A, ray, R.steps);
    drop_state(A.st, i);
}

// Claims work items until one maps to a ray still to trace or the queue is
// exhausted; `slot` is the item's index within the launch.  RESUME: the rays
// were saved by the approach kernel.
template <bool RESUME>
__device__ inline bool claim(const Params &P, const TraceArgs &A, int &ray, int &slot)
{
    while (true) {
        const int q = (int)atomicAdd(A.queue, 1u);
        if (q >= A.total)
            return false;
        slot = q;
        if (RESUME && q < A.nsaved) {
            ray = state_ray(A.st, q);
            if (ray >= 0)
                return true;
        } else if (work_to_ray(P, A, A.w0 + q, ray)) {
            return true;
        }
    }
}

template <bool RESUME, int NF, bool GRAD>
__device__ inline void start_ray(const Params &P, const TraceArgs &A, int ray, int slot,
                                 Ray<NF, GRAD> &R)
{
    if (RESUME && slot < A.nsaved)
        load_state(A.st, slot, R);
    else
        ray_init<NF, GRAD>(P, ray / P.npix, ray % P.npix, R);
}

// Each lane traces one ray at a time.  A lane whose ray has ended waits until
// at least `refill` lanes of its wave are idle; then all idle lanes claim new
// ray ids together.  Lanes claiming together get consecutive ids (the atomic
// is combined across the wave), i.e. neighbouring pixels, whose rays stay in
// step for most of their length; refill = 1 refills each lane at once,
// refill = 64 refills whole waves.  Waiting for a few idle lanes (default 4)
// runs the ray set-up once for several lanes at the cost of a few idle lane
// steps.  The loop ends when the queue is exhausted and every lane is idle,
// so every wave drains.
// One frequency (except the 3D gradient pass) fits in 168 VGPRs without
// spilling, so those kernels are also built for 3 waves per SIMD, which hide
// more of the FP64 dependency latency: 6% on the 1024^2 image.  Images of up
// to 8 frequencies of a 2D dump without gradients are built so too: at 8
// frequencies the bound spills a few per-ray counters, reloaded once a step,
// and the third wave still gains 10% (6 inclinations x 8 frequencies at
// 512^2).  A launch
// with few rays per lane is instead bound by its longest rays, which a third
// wave sharing the SIMD slows down (512^2: 10% slower), so the 3-wave build
// is used from kThreeWaveRays rays per launch.  The other kernels are left to
// the compiler, which fits them in 2 waves where it can: a forced bound only
// trades its instruction scheduling for registers they do not need
// (measured slower).
static constexpr long long kThreeWaveRays = 1LL << 19;
static constexpr int kSimdsPerCu = 4;
static constexpr int kLongRay = 800;   // steps: about the 99.5th percentile at 1024^2
template <int NF, bool GRAD, bool DIM3, int INT>
constexpr bool three_waves()
{
    return INT == kRK2 && (NF == 1 ? !(GRAD && DIM3) : (NF <= 8 && !GRAD && !DIM3));
}

template <int NF, bool GRAD, bool DIM3, int INT, bool RESUME, int EMIS>
__device__ __forceinline__ void trace_rays(const Params &P, const TraceArgs &A)
{
    // resumed RK2 rays overlap each transfer sample with the next step's
    // first acceleration (ray_step_rk2_overlap), carried in A0
    constexpr bool kOverlap = RESUME && INT == kRK2;
    // a ray whose FP32 optical depth met the ln 1000 gate where a different
    // decision would show in the pixel (needs_exact): with one frequency of
    // the thermal fit, in the kernels tracing from the camera, its lane
    // traces it again at once with the FP64 transfer (xm); the resuming
    // kernels (whose registers that branch would spill: 1024^2 render -6%
    // without it) and those of several frequencies list it for exact_kernel
    constexpr bool kInline = EMIS == kThSynch && NF == 1 && !RESUME;
    Ray<NF, GRAD> R;
    double A0[4] = {0., 0., 0., 0.};
    int ray = 0, slot = 0;
    bool xm = false;
    bool active = claim<RESUME>(P, A, ray, slot);
    bool exhausted = !active;
    if (active) {
        start_ray<RESUME>(P, A, ray, slot, R);
        if (kOverlap)
            geodesic_accel(R.pt, P.a, P.hslope, R.y + 4, A0);
    }
    while (true) {
        if (active) {
            bool cont = ray_alive<NF, GRAD>(P, R);
            if (cont) {
                if constexpr (kInline && kOverlap)
                    cont = ray_step_rk2_overlap_mixed<NF, GRAD, DIM3>(P, A.prims, R, A0, xm);
                else if constexpr (kInline)
                    cont = ray_step_mixed<NF, GRAD, DIM3, INT, !RESUME>(P, A.prims, R, xm);
                else if constexpr (kOverlap)
                    cont = ray_step_rk2_overlap<NF, GRAD, DIM3, EMIS>(P, A.prims, R, A0);
                else
                    cont = ray_step<NF, GRAD, DIM3, INT, EMIS, false, !RESUME>(P, A.prims, R);
            }
            if (!cont) {
                if (kInline && !xm && needs_exact<EMIS>(R)) {
                    start_ray<RESUME>(P, A, ray, slot, R);
                    if (kOverlap)
                        geodesic_accel(R.pt, P.a, P.hslope, R.y + 4, A0);
                    xm = true;
                } else {
                    write_ray<NF, GRAD, DIM3>(P, A, ray, R);
                    if (!kInline && needs_exact<EMIS>(R))
                        A.xlist[atomicAdd(A.xcount, 1u)] = make_int2(slot, (int)R.acc.gmask);
                    active = false;
                    xm = false;
                }
            }
        }
        // A wave holding a ray that has run kLongRay steps (a ray near the
        // photon ring: these orbit the hole and take up to a few thousand
        // steps) issues ahead of the other waves on its SIMD, so that the
        // launch does not end waiting on it.  Kernels tracing from the
        // camera only: the resuming kernels' longest rays are the lead
        // kernel's, and there the two ballots a step cost more (1024^2
        // render -2%).
        if constexpr (!RESUME) {
            if (__ballot(active && R.steps > kLongRay))
                __builtin_amdgcn_s_setprio(2);
            else
                __builtin_amdgcn_s_setprio(0);
        }
        const unsigned long long act = __ballot(active);
        exhausted = __ballot(exhausted) != 0ull;
        const int idle = 64 - __popcll(act);
        if (exhausted) {
            if (act == 0ull)
                break;
        } else if (idle >= A.refill || (act == 0ull)) {
            if (!active) {
                if (claim<RESUME>(P, A, ray, slot)) {
                    start_ray<RESUME>(P, A, ray, slot, R);
                    if (kOverlap)
                        geodesic_accel(R.pt, P.a, P.hslope, R.y + 4, A0);
                    active = true;
                } else {
                    exhausted = true;
                }
            }
        }
    }
}

template <int NF, bool GRAD, bool DIM3, int INT, bool RESUME, int EMIS>
__global__ __launch_bounds__(64) void trace_kernel(const Params P, const TraceArgs A)
{
    trace_rays<NF, GRAD, DIM3, INT, RESUME, EMIS>(P, A);
}

template <int NF, bool GRAD, bool DIM3, int INT, bool RESUME, int EMIS>
__global__ __launch_bounds__(64, 3) void trace_kernel_3w(const Params P, const TraceArgs A)
{
    trace_rays<NF, GRAD, DIM3, INT, RESUME, EMIS>(P, A);
}

// Lane k's value of v (k the same on every lane).
__device__ inline double lane_read(double v, int k)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, k);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), k);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// exact_retrace<1, GRAD, DIM3, INT, false> (the FP64 transfer alone, one
// frequency) by a whole wave: every lane integrates the same geodesic, 64
// steps at a time, lane k keeping the point of the batch's step k; the lanes
// then evaluate their points' FP64 samples at once (sample_terms_f64), and
// the optical depth, intensity and tangents are summed over the samples in
// step order as the one-lane re-trace sums them, up to the sample after
// which the frequency is opaque.  The same arithmetic, so the same bits; the
// batch's geodesic steps after that sample are dropped.  One lane's re-trace
// costs a wave's issue per step either way; here the FP64 sample, most of a
// step, is shared by the batch.
template <bool GRAD, bool DIM3, int INT>
__device__ void exact_retrace_wave(const Params &P, const float4 *__restrict__ prims,
                                   Ray<1, GRAD> &R, Accum<1, GRAD> &x, int f)
{
    const int lane = (int)threadIdx.x;
    const double lnth = 6.907755278982137;   // log(1000.)
    x.clear();
    bool done = false;
    while (!done) {
        double my[8];
        Point mpt = R.pt;
        double mdl = 0.;
        bool min = false;
        int nb = 0;
#pragma unroll 1
        for (int k = 0; k < 64; k++) {
            double dl;
            if (!ray_alive<1, GRAD>(P, R) || !ray_advance<INT>(P, R, dl)) {
                done = true;
                break;
            }
            if (lane == k) {
#pragma unroll
                for (int q = 0; q < 8; q++)
                    my[q] = R.y[q];
                mpt = R.pt;
                mdl = dl;
                min = in_grid(P, R.y);
            }
            R.steps++;
            nb++;
        }
        SampleF64 T = {};
        bool ok = false;
        if (lane < nb && min) {
            Fluid64 F;
            if (fluid_sample_f64<DIM3>(P, prims, mpt, my, F)) {
                ok = true;
                T = sample_terms_f64<GRAD>(F, synch_f64(F), P.freq[f], fabs(mdl) * P.Rg);
            }
        }
        // optical depth after each sample, in step order, to the sample that
        // makes the frequency opaque (the one-lane re-trace stops there)
        const unsigned long long okm = __ballot(ok);
        double tau = tau_f64(x.tau[0], x.ctau[0]), mtau = 0.;
        int last = 64;
        for (unsigned long long m = okm; m; m &= m - 1) {
            const int k = __builtin_ctzll(m);
            tau = tau + lane_read(T.dtau, k);
            if (lane == k)
                mtau = tau;
            if (!(tau < lnth)) {
                last = k;
                break;
            }
        }
        const double metau = dexp(-mtau);
        for (unsigned long long m = okm; m; m &= m - 1) {
            const int k = __builtin_ctzll(m);
            if (k > last)
                break;
            SampleF64 t;
            t.dtau = lane_read(T.dtau, k);
            t.j = lane_read(T.j, k);
            t.dl = lane_read(T.dl, k);
            t.gi2 = lane_read(T.gi2, k);
            if (GRAD) {
                t.dlnj0 = lane_read(T.dlnj0, k);
                t.dlnj1 = lane_read(T.dlnj1, k);
                t.dlna1 = lane_read(T.dlna1, k);
            }
            const double tk = lane_read(mtau, k);
            set_tau_f64(x.tau[0], x.ctau[0], tk);
            add_sample_f64(x, 0, t, tk, lane_read(metau, k));
        }
        // the ray's counters as the one-lane re-trace leaves them
        if (last < 64) {
            R.steps -= nb - 1 - last;
            R.samples += __popcll(okm & ((2ull << last) - 1ull));
            done = true;
        } else {
            R.samples += __popcll(okm);
        }
    }
}

// The rays a trace kernel listed (needs_exact) traced again from their start
// (the camera, or the saved ray at the grid's edge) with the FP64 transfer.
// Rows blockIdx.y < n_freq, one frequency per lane: row f takes frequency f
// of every listed ray whose gate mask has bit f (the frequencies that had a
// sample in the gate window; for the others the FP32 outputs stand) and
// several bits; where that frequency's own share of the window exceeds
// kGateFlip it gets the FP64 transfer's outputs, written over its FP32 ones
// -- the same decision and the same arithmetic as a launch of that
// frequency alone, which traces the ray again in its trace kernel from the
// camera, or here.  A ray with one bit was listed for that frequency's own
// share, so the decision is known and only the FP64 transfer runs, by a
// whole wave in the rows past n_freq (exact_retrace_wave).  A second trace
// ends where its frequency's optical depth has passed the gate, so that it
// runs from the grid's edge to the crossing, not to the end of the ray at
// every frequency.  A kernel of its own: its FP64 arithmetic needs more
// registers than the resuming and multi-frequency trace kernels can give.
// Consecutive entries go to different waves, so that those few rays run on
// different SIMDs.
template <bool GRAD, bool DIM3, int INT, bool RESUME>
__global__ __launch_bounds__(64) void exact_kernel(const Params P, const TraceArgs A)
{
    const unsigned n = *A.xcount;
    if ((int)blockIdx.y >= P.nf) {
        // the rows past the frequencies: a wave per listed ray with one
        // frequency in its gate mask (exact_retrace_wave)
        const unsigned rows = gridDim.y - (unsigned)P.nf;
        const unsigned stride = gridDim.x * rows;
        for (unsigned i = blockIdx.x + gridDim.x * (blockIdx.y - (unsigned)P.nf); i < n;
             i += stride) {
            const int2 e = A.xlist[i];
            const unsigned mask = (unsigned)e.y;
            if (mask == 0u || (mask & (mask - 1u)) != 0u)
                continue;
            const int f = __builtin_ctz(mask);
            const int q = e.x;
            int ray;
            if (RESUME && q < A.nsaved)
                ray = state_ray(A.st, q);
            else if (!work_to_ray(P, A, A.w0 + q, ray))
                continue;
            Ray<1, GRAD> R;
            start_ray<RESUME>(P, A, ray, q, R);
            Accum<1, GRAD> x;
            exact_retrace_wave<GRAD, DIM3, INT>(P, A.prims, R, x, f);
            if (threadIdx.x == 0) {
                const size_t o = (size_t)ray * P.nf + f;
                A.image[o] = (double)x.I[0];
                if (GRAD) {
                    A.dimage[o * 2 + 0] = x.dI[0][0];
                    A.dimage[o * 2 + 1] = x.dI[0][1];
                }
                // one frequency: the counters of the re-trace, as a trace
                // kernel that traces its rays again itself writes them
                if (A.counters && P.nf == 1) {
                    rt_ray_counters c;
                    c.steps = R.steps;
                    c.samples = R.samples;
                    A.counters[ray] = c;
                }
            }
        }
        return;
    }
    const int f = blockIdx.y;
    const unsigned stride = gridDim.x * 64u;
    for (unsigned i = blockIdx.x + gridDim.x * threadIdx.x; i < n; i += stride) {
        const int2 e = A.xlist[i];
        // a ray listed with one frequency in the gate window was listed for
        // that frequency's own share (the first trace summed no other), which
        // is the decision: the wave rows trace its FP64 transfer again
        if (!((e.y >> f) & 1) || (e.y & (e.y - 1)) == 0)
            continue;
        const int q = e.x;
        int ray;
        if (RESUME && q < A.nsaved)
            ray = state_ray(A.st, q);
        else if (!work_to_ray(P, A, A.w0 + q, ray))
            continue;
        Ray<1, GRAD> R;
        start_ray<RESUME>(P, A, ray, q, R);
        Accum<1, GRAD> x;
        float wf[1];
        exact_retrace<1, GRAD, DIM3, INT, true, RESUME && INT == kRK2>(P, A.prims, R, x, wf, f);
        if (wf[0] > kGateFlip) {
            const size_t o = (size_t)ray * P.nf + f;
            A.image[o] = (double)x.I[0];
            if (GRAD) {
                A.dimage[o * 2 + 0] = x.dI[0][0];
                A.dimage[o * 2 + 1] = x.dI[0][1];
            }
        }
    }
}

// The energy spectrum (core.c:146-150): per camera and frequency the sum of
// the image over its pixels, in two passes of fixed order -- partial sums of
// kSpecBlocks contiguous pixel ranges, then their sum -- so that every call
// gives the same bits.  Thread t of a block sums every 256th pixel of the
// block's range from the t-th on, at each frequency (NF >= nf accumulators),
// then a tree over the block's threads adds those up per frequency: the
// order of each (camera, frequency) sum does not depend on how many
// frequencies the image has, so batched and single-frequency renders give
// the same spectrum.
static constexpr int kSpecBlocks = 256;

template <int NF>
__global__ __launch_bounds__(256) void spectrum_partial(const double *img, int npix, int nf,
                                                        double *part)
{
    __shared__ double red[256];
    const int cam = blockIdx.y, nb = gridDim.x;
    const int per = (npix + nb - 1) / nb;
    const int p0 = blockIdx.x * per, p1 = min(npix, p0 + per);
    const int t = threadIdx.x;
    const double *im = img + (size_t)cam * npix * nf;
    double sum[NF];
#pragma unroll
    for (int f = 0; f < NF; f++)
        sum[f] = 0.;
    for (int i = p0 + t; i < p1; i += 256) {
#pragma unroll
        for (int f = 0; f < NF; f++)
            if (f < nf)
                sum[f] += im[(size_t)i * nf + f];
    }
#pragma unroll
    for (int f = 0; f < NF; f++) {
        if (f < nf) {   // the same in every thread
            red[t] = sum[f];
            __syncthreads();
            for (int w = 128; w > 0; w >>= 1) {
                if (t < w)
                    red[t] += red[t + w];
                __syncthreads();
            }
            if (t == 0)
                part[((size_t)cam * nb + blockIdx.x) * nf + f] = red[0];
            __syncthreads();
        }
    }
}

__global__ void spectrum_final(const double *part, int nb, int nf, double *spec)
{
    const int cam = blockIdx.x, f = threadIdx.x;
    if (f >= nf)
        return;
    double s = 0.;
    for (int b = 0; b < nb; b++)
        s += part[((size_t)cam * nb + b) * nf + f];
    spec[cam * nf + f] = s;
}

// q[0]: the trace kernel's claim counter, q[1]: the length of its exact list
__global__ void reset_queue_kernel(unsigned int *q)
{
    q[0] = 0u;
    q[1] = 0u;
}

// ---- probes: one thread per point ----------------------------------------
__global__ void probe_metric(const Params P, const double *in, int n, double *out)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const double *X = in + 4 * t;
    Point pt = make_point(X[1], X[2], P.hslope);
    MetricDD g = metric_dd(pt, P.a, P.hslope);
    double *o = out + 32 * t;
    for (int i = 0; i < 32; i++)
        o[i] = 0.;
    o[0] = g.g00;
    o[1] = o[4] = g.g01;
    o[3] = o[12] = g.g03;
    o[5] = g.g11;
    o[7] = o[13] = g.g13;
    o[10] = g.g22;
    o[15] = g.g33;
    double r = pt.r, a = P.a;
    double irho2 = 1. / (r * r + a * a * pt.c * pt.c);
    double hfac = kPi + (1. - P.hslope) * kPi * pt.c2;
    double *u = o + 16;
    u[0] = -1. - 2. * r * irho2;
    u[1] = u[4] = 2. * irho2;
    u[5] = irho2 * (r * (r - 2.) + a * a) / (r * r);
    u[7] = u[13] = a * irho2 / r;
    u[10] = irho2 / (hfac * hfac);
    u[15] = irho2 / (pt.s * pt.s);
}

__global__ void probe_accel(const Params P, const double *in, int n, double *out)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const double *X = in + 8 * t;
    Point pt = make_point(X[1], X[2], P.hslope);
    geodesic_accel(pt, P.a, P.hslope, X + 4, out + 4 * t);
}

__global__ void probe_rk2(const Params P, const double *in, int n, double *out)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    double y[8];
    for (int i = 0; i < 8; i++)
        y[i] = in[8 * t + i];
    Point pt = make_point(y[1], y[2], P.hslope);
    double dl = stepsize(y, P.stepsize);
    rk2_step(y, pt, dl, P.a, P.hslope);
    out[9 * t] = dl;
    for (int i = 0; i < 8; i++)
        out[9 * t + 1 + i] = y[i];
}

__global__ void probe_camera(const Params P, const double *in, int n, double *out)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    int cam = (int)in[3 * t + 2];
    init_photon(P.cam[cam], in[3 * t], in[3 * t + 1], out + 8 * t);
}

__global__ void probe_emission_model(const KappaConst K, const double *in, int n, double *out)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const double *q = in + 6 * t;
    const int model = (int)q[5];
    const float B = (float)q[0], mu = (float)q[1], Te = (float)q[2], nu = (float)q[3],
                ne = (float)q[4];
    if (model == kThSynch) {
        EmisF e = emission_f32(B, mu, Te, nu, ne);
        double jx = (double)e.jh * exp(-(double)e.cb);
        out[2 * t] = kE2C * jx;
        out[2 * t + 1] = e.jh != 0.f ? jx * (double)e.kap : (e.poison ? NAN : 0.);
        return;
    }
    const EmisX e = model == kKappa ? emission_kappa(K, B, mu, Te, nu, ne)
                                    : emission_thsynchav(B, Te, nu, ne);
    out[2 * t] = kE2C * (double)e.jx;
    out[2 * t + 1] = (model == kKappa || e.jx != 0.f) ? (double)e.alpha : (e.poison ? NAN : 0.);
}

__global__ void probe_emission(const double *in, int n, double *out)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const double *q = in + 5 * t;
    EmisF e = emission_f32((float)q[0], (float)q[1], (float)q[2], (float)q[3], (float)q[4]);
    double jx = (double)e.jh * exp(-(double)e.cb);
    out[2 * t] = kE2C * jx;
    out[2 * t + 1] = e.jh != 0.f ? jx * (double)e.kap : (e.poison ? NAN : 0.);
}

template <bool DIM3>
__global__ void probe_fluid(const Params P, const float4 *prims, const double *in, int n,
                            double *out)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    double y[8];
    for (int i = 0; i < 4; i++) {
        y[i] = in[4 * t + i];
        y[4 + i] = 0.;
    }
    y[4] = 1.;
    Point pt = make_point(y[1], y[2], P.hslope);
    Fluid F;
    float u[4], b[4], bb[2];
    bool ok = fluid_sample<DIM3>(P, prims, pt, y, F, u, b, bb);
    double *o = out + 14 * t;
    o[0] = ok ? F.ne : kSmall;
    o[1] = ok ? F.thetae : kSmall;
    o[2] = ok ? F.B : kSmall;
    for (int i = 0; i < 4; i++) {
        o[3 + i] = u[i];
        o[7 + i] = b[i];
    }
    o[11] = ok ? 1. : 0.;
    o[12] = bb[0];
    o[13] = bb[1];
}

// cell-interleaved staging: 8 primitives of one cell in two float4
__global__ void stage_prims(const float *soa, size_t ncell, float4 *aos)
{
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell)
        return;
    aos[2 * c] = make_float4(soa[c], soa[ncell + c], soa[2 * ncell + c], soa[3 * ncell + c]);
    aos[2 * c + 1] = make_float4(soa[4 * ncell + c], soa[5 * ncell + c], soa[6 * ncell + c],
                                 soa[7 * ncell + c]);
}

}  // namespace

// claim-order table key: W, H, tile, pixel size bits (x, y), and for the
// ring order the spin and inclination bits (0 otherwise)
static constexpr int kOrderKey = 7;

struct rt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // boundaries between a one-chunk render's kernels: after the approach,
    // after the trace kernel, after the FP64 re-trace (rt_last_kernel_times)
    hipEvent_t evA = nullptr, evT = nullptr, evX = nullptr;
    // the ring order's lead launch: a second stream (default priority),
    // forked from and joined back into `stream` by the two events, and the
    // lead kernel's own time (evL0, evL1; lead_timed: they hold the last
    // render's)
    hipStream_t lead_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, evL0 = nullptr, evL1 = nullptr;
    bool lead_timed = false;
    int last_chunks = 0;   // launch chunks of the last eager render
    int num_cu = 0;
    bool have_params = false, have_grid = false;
    rt_params params{};
    rt_grid grid{};
    float4 *d_prims = nullptr;
    // Live graphs that launch with d_prims.  rt_load_prims keeps a buffer
    // that graphs still use in retired_prims (they go on rendering the dump
    // they were captured with) until the last of them is destroyed.
    int prims_refs = 0;
    struct RetiredPrims {
        void *d;
        int refs;
    };
    std::vector<RetiredPrims> retired_prims;
    unsigned int *d_queue = nullptr;
    // Centre-out claim orders, one per image shape (W, H, tile, pixel size
    // bits); beyond kMaxOrderTables the least recently used one is freed.
    // A table is never changed in place, and one that a captured graph
    // launches with is pinned until that graph is destroyed.
    struct OrderTable {
        long long key[kOrderKey];
        int *d;
        int nlead;                 // ring order: its first nlead entries are the lead rays
        int refs;                  // live graphs that use the table
        unsigned long long used;   // last use, for eviction
    };
    std::vector<OrderTable> orders;
    unsigned long long use_clock = 0;
    // Saved-ray buffers of the split launch.  back() is the current one; an
    // older one stays only while a captured graph launches with it.
    struct StateBuf {
        void *d;
        int slots;
        int refs;
    };
    std::vector<StateBuf> states;
    std::vector<void *> capture_pins;   // tables and buffers used by the open capture
    // work items of one launch to trace again with the FP64 transfer
    // (exact_kernel), kMaxChunk of them: made by the first eager render and
    // kept, so that a captured graph's launches always find it
    int2 *d_xlist = nullptr;
    // spectrum: partial sums [camera][kSpecBlocks][frequency] and, for host
    // renders, the spectrum itself; made by rt_create
    double *d_spec_part = nullptr, *d_spec = nullptr;
    // this device's `scratch` (rt_render_device_noalloc), from
    // hipGetSymbolAddress at rt_create
    void *d_scratch = nullptr;
    double *d_image = nullptr, *d_dimage = nullptr;
    size_t image_cap = 0, dimage_cap = 0;
    // the kappa absorptivity's hypergeometric table (kappa_table), made once
    float2 *d_ktab = nullptr;
    float ktab_u0 = 0.f, ktab_idu = 1.f;
    int ktab_n = 0;
    double last_ms = 0.;
    rt_launch_config cfg{64, 0, kDefaultRefill, kDefaultTile, RT_ORDER_RING, 0};
    bool capturing = false;   // between rt_capture_begin and rt_capture_end
    char err[512] = {0};
};

// The context's kappa table in the kappa fit's constants.
static KappaConst kappa_with_table(const rt_ctx *c)
{
    KappaConst K = kappa_const();
    K.tab = c->d_ktab;
    K.tu0 = c->ktab_u0;
    K.tidu = c->ktab_idu;
    K.tn = c->ktab_n;
    return K;
}

struct rt_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int device = 0;
    std::vector<void *> pins;   // order tables and state buffers its launches use
};

static constexpr size_t kMaxOrderTables = 8;
// Work items per approach + trace launch pair with the split launch: 2^23
// saved rays are 1 GB; larger batches run in chunks of this size.
static constexpr int kMaxChunk = 1 << 23;

// rt_render_device_noalloc's device memory: one static array per GPU, held
// until the process ends, in place of the buffers rt_render_device
// allocates -- the work queue's two counters, the FP64 re-trace's list (one
// int per work item) and the spectrum's partial sums.  Every call on a GPU
// shares its array: calls that overlap on one GPU race.  Host code uses the
// address hipGetSymbolAddress returns (scratch_address), never the name.
static constexpr int kNoAllocRays = RT_NOALLOC_MAX_RAYS;
static constexpr size_t kScratchQueue = 0;     // 2 unsigned ints, 256 bytes reserved
static constexpr size_t kScratchXlist = 256;
static constexpr size_t kScratchSpec = kScratchXlist + (size_t)kNoAllocRays * sizeof(int2);
static constexpr size_t kScratchBytes =
    kScratchSpec + (size_t)kMaxCam * kSpecBlocks * kMaxFreq * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch[kMaxScratchDevices];

// The device address of `scratch` on `device`, which must be current:
// resolved once per device.
static hipError_t scratch_address(int device, void **out)
{
    if (device < 0 || device >= kMaxScratchDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        if (!p)
            return hipErrorInvalidSymbol;
        g_scratch[device] = p;
    }
    *out = g_scratch[device];
    return hipSuccess;
}

static void pin(rt_ctx *c, void *p)
{
    if (std::find(c->capture_pins.begin(), c->capture_pins.end(), p) == c->capture_pins.end())
        c->capture_pins.push_back(p);
}

static char g_create_err[512];

static int fail(rt_ctx *c, int code, const char *fmt, ...)
{
    char *buf = c ? c->err : g_create_err;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, 512, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(ctx, call)                                                              \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(ctx, RT_E_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                           \
    } while (0)

static int select_device(rt_ctx *c)
{
    HIPCHK(c, hipSetDevice(c->device));
    return RT_OK;
}

// Fills the launch parameters from the context; validates sizes.
static int make_params(rt_ctx *c, const double *incl, int ncam, const double *freqs, int nf,
                       Params &P)
{
    if (!c->have_params)
        return fail(c, RT_E_STATE, "rt_set_params has not been called");
    if (!c->have_grid)
        return fail(c, RT_E_STATE, "rt_load_prims has not been called");
    if (ncam < 1 || ncam > kMaxCam)
        return fail(c, RT_E_INVALID, "n_cam must be in [1, %d], got %d", kMaxCam, ncam);
    if (nf < 1 || nf > kMaxFreq)
        return fail(c, RT_E_INVALID, "n_freq must be in [1, %d], got %d", kMaxFreq, nf);
    if (!incl || !freqs)
        return fail(c, RT_E_INVALID, "inclinations and freqs must not be NULL");
    const rt_params &p = c->params;
    const rt_grid &g = c->grid;
    long long total = (long long)ncam * p.width * p.height;
    if (total > 0x7fffff00LL)
        return fail(c, RT_E_INVALID, "too many rays in one launch (%lld)", total);
    const char *msg = fill_params(p, g, incl, ncam, freqs, nf, P);
    if (msg)
        return fail(c, RT_E_INVALID, "%s", msg);
    P.kc = kappa_with_table(c);
    return RT_OK;
}

template <int NF, bool GRAD, bool DIM3, int INT, bool RESUME, int EMIS = kThSynch>
static hipError_t launch_trace(rt_ctx *c, const Params &P, const TraceArgs &A,
                               hipStream_t stream)
{
    void (*kern)(const Params, const TraceArgs) = trace_kernel<NF, GRAD, DIM3, INT, RESUME, EMIS>;
    if constexpr (EMIS == kThSynch && three_waves<NF, GRAD, DIM3, INT>()) {
        if (A.total >= kThreeWaveRays)
            kern = trace_kernel_3w<NF, GRAD, DIM3, INT, RESUME, EMIS>;
    }
    int block = 64;
    int per_cu = c->cfg.blocks_per_cu;
    if (per_cu <= 0) {
        int occ = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, block, 0);
        if (e != hipSuccess)
            return e;
        per_cu = std::max(occ, 1);
        // A launch of a few rays per lane ends on its longest rays (photon-ring
        // rays of thousands of steps), whose waves then run at the issue rate
        // they get on their SIMD: one wave per SIMD (a CU's 4 SIMDs, one
        // 64-lane workgroup each) instead of two, 7-8% faster at 512^2 (2D
        // image, 3D image and gradient).
        if (A.total < kThreeWaveRays)
            per_cu = std::min(per_cu, kSimdsPerCu);
    }
    long long grid = (long long)per_cu * c->num_cu;
    long long need = ((long long)A.total + 63) / 64;
    grid = std::max(1LL, std::min(grid, need));
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(block), 0, stream, P, A);
    return hipGetLastError();
}

// The thermal fit is built for 1, 4, 8 and 16 frequencies, image and
// gradient; the other emission models (image only) for 1 and 16.
template <bool GRAD, bool DIM3, int INT, bool RESUME>
static hipError_t dispatch_nf(rt_ctx *c, const Params &P, const TraceArgs &A,
                              hipStream_t st)
{
    if (P.emission != kThSynch) {
        if constexpr (GRAD) {
            return hipErrorInvalidValue;   // refused before the launch
        } else {
            if (P.emission == kKappa)
                return P.nf == 1 ? launch_trace<1, false, DIM3, INT, RESUME, kKappa>(c, P, A, st)
                                 : launch_trace<16, false, DIM3, INT, RESUME, kKappa>(c, P, A, st);
            return P.nf == 1 ? launch_trace<1, false, DIM3, INT, RESUME, kThSynchAv>(c, P, A, st)
                             : launch_trace<16, false, DIM3, INT, RESUME, kThSynchAv>(c, P, A, st);
        }
    }
    if (P.nf == 1)
        return launch_trace<1, GRAD, DIM3, INT, RESUME>(c, P, A, st);
    if (P.nf <= 4)
        return launch_trace<4, GRAD, DIM3, INT, RESUME>(c, P, A, st);
    if (P.nf <= 8)
        return launch_trace<8, GRAD, DIM3, INT, RESUME>(c, P, A, st);
    return launch_trace<16, GRAD, DIM3, INT, RESUME>(c, P, A, st);
}

// The FP64 re-trace after a resuming or multi-frequency trace kernel
// (thermal fit only; a one-frequency kernel tracing from the camera traces
// its rays again itself): n_freq rows of up to 4 workgroups per CU, each lane
// taking list entries in a row-sized stride.
template <bool GRAD, bool DIM3, int INT, bool RESUME>
static hipError_t dispatch_exact(rt_ctx *c, const Params &P, const TraceArgs &A)
{
    if (P.emission != kThSynch || (P.nf == 1 && !RESUME))
        return hipSuccess;
    const long long gx = std::max(1LL, std::min<long long>(4LL * c->num_cu,
                                                           ((long long)A.total + 63) / 64));
    // and rows of waves for the rays with one frequency in the window, about
    // 8 waves per CU
    const long long wrows = std::max(1LL, (8LL * c->num_cu + gx - 1) / gx);
    hipLaunchKernelGGL((exact_kernel<GRAD, DIM3, INT, RESUME>),
                       dim3((unsigned)gx, (unsigned)(P.nf + wrows)),
                       dim3(64), 0, c->stream, P, A);
    return hipGetLastError();
}

// Work items chunk by chunk (at most kMaxChunk each): with pct <= 0 all in
// one persistent launch per chunk; otherwise the approach kernel over the
// first pct percent of the chunk's work items, then the persistent kernel,
// which resumes those and traces the rest from the camera (a chunk saves at
// most A.st.n rays).  After each, the rays listed for the FP64 transfer.
template <bool GRAD, bool DIM3, int INT>
static hipError_t run_launches(rt_ctx *c, const Params &P, TraceArgs A, int total, int pct,
                               const TraceArgs *lead)
{
    hipError_t e;
    if (lead) {
        // the ring's rays from the camera on the lead stream, from the start
        // of the render; joined back after the chunk's kernels below (one
        // chunk: the ring order is used for launches of at most kMaxChunk)
        const bool timed = !c->capturing;
        if ((e = hipEventRecord(c->ev_fork, c->stream)) != hipSuccess ||
            (e = hipStreamWaitEvent(c->lead_stream, c->ev_fork, 0)) != hipSuccess)
            return e;
        if (timed && (e = hipEventRecord(c->evL0, c->lead_stream)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, c->lead_stream, lead->queue);
        if ((e = hipGetLastError()) != hipSuccess ||
            (e = dispatch_nf<GRAD, DIM3, INT, false>(c, P, *lead, c->lead_stream)) != hipSuccess)
            return e;
        if (timed && (e = hipEventRecord(c->evL1, c->lead_stream)) != hipSuccess)
            return e;
        if ((e = hipEventRecord(c->ev_join, c->lead_stream)) != hipSuccess)
            return e;
        c->lead_timed = timed;
    } else if (!c->capturing) {
        c->lead_timed = false;
    }
    const long long chunk =
        pct <= 0 ? (long long)kMaxChunk
                 : std::min<long long>(kMaxChunk, (long long)A.st.n * 100 / pct);   // saves <= st.n
    // a one-chunk eager render times its kernels one by one
    const bool mark = !c->capturing && total <= chunk;
    if (!c->capturing)
        c->last_chunks = (int)((total + chunk - 1) / chunk);
    for (long long w0 = 0; w0 < total; w0 += chunk) {
        A.w0 = (int)w0;
        A.total = (int)std::min<long long>(chunk, total - w0);
        A.nsaved = pct <= 0 ? 0 : (int)(((long long)A.total * pct + 99) / 100);
        if (A.nsaved > 0)
            hipLaunchKernelGGL(approach_kernel<INT>, dim3((unsigned)((A.nsaved + 63) / 64)),
                               dim3(64), 0, c->stream, P, A);
        hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, c->stream, A.queue);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
        if (mark && (e = hipEventRecord(c->evA, c->stream)) != hipSuccess)
            return e;
        e = A.nsaved > 0 ? dispatch_nf<GRAD, DIM3, INT, true>(c, P, A, c->stream)
                         : dispatch_nf<GRAD, DIM3, INT, false>(c, P, A, c->stream);
        if (e != hipSuccess)
            return e;
        if (mark && (e = hipEventRecord(c->evT, c->stream)) != hipSuccess)
            return e;
        e = A.nsaved > 0 ? dispatch_exact<GRAD, DIM3, INT, true>(c, P, A)
                         : dispatch_exact<GRAD, DIM3, INT, false>(c, P, A);
        if (e != hipSuccess)
            return e;
        if (mark && (e = hipEventRecord(c->evX, c->stream)) != hipSuccess)
            return e;
    }
    if (lead && (e = hipStreamWaitEvent(c->stream, c->ev_join, 0)) != hipSuccess)
        return e;
    return hipSuccess;
}

template <bool GRAD, bool DIM3>
static hipError_t dispatch_int(rt_ctx *c, const Params &P, const TraceArgs &A, int total, int pct,
                               const TraceArgs *lead)
{
    if (c->params.integrator == RT_INT_RK4)
        return run_launches<GRAD, DIM3, kRK4>(c, P, A, total, pct, lead);
    if (c->params.integrator == RT_INT_VERLET)
        return run_launches<GRAD, DIM3, kVerlet>(c, P, A, total, pct, lead);
    return run_launches<GRAD, DIM3, kRK2>(c, P, A, total, pct, lead);
}

static int ensure_buffer(rt_ctx *c, double **buf, size_t *cap, size_t bytes)
{
    if (*cap >= bytes)
        return RT_OK;
    if (*buf)
        HIPCHK(c, hipFree(*buf));
    *buf = nullptr;
    *cap = 0;
    HIPCHK(c, hipMalloc((void **)buf, bytes));
    *cap = bytes;
    return RT_OK;
}

// The saved-ray buffer, at least `slots` rays (rounded up to 64).  A capture
// uses the current buffer as it is (in more, smaller chunks if it is short)
// and is refused if there is none.
static int ensure_state(rt_ctx *c, long long slots, RayState *out)
{
    const int want = (int)((slots + 63) / 64 * 64);
    if (c->states.empty() || (c->states.back().slots < want && !c->capturing)) {
        if (c->capturing)
            return fail(c, RT_E_STATE, "the split launch's ray buffer is allocated by an eager "
                                       "render; render once before capturing");
        if (!c->states.empty() && c->states.back().refs == 0) {
            HIPCHK(c, hipStreamSynchronize(c->stream));   // launches still reading it
            HIPCHK(c, hipFree(c->states.back().d));
            c->states.pop_back();
        }
        rt_ctx::StateBuf b{nullptr, want, 0};
        HIPCHK(c, hipMalloc(&b.d, (size_t)want * kStateBytes));
        c->states.push_back(b);
    }
    const rt_ctx::StateBuf &b = c->states.back();
    if (c->capturing)
        pin(c, b.d);
    out->d = (char *)b.d;
    out->n = b.slots;
    return RT_OK;
}

// Centre-out claim order: the T x T tiles sorted by the distance of their
// centre from the image centre (in camera units, so non-square pixels are
// measured correctly), ties in row-major tile order, pixels row-major inside
// a tile.  Rays near the centre -- the shadow and the photon ring, up to ten
// times longer than escaping rays -- are claimed first, so the launch does
// not end waiting on a few long rays that started late.  Built on the host
// once per image size and tile, kept in HBM (4 bytes per pixel).
//
// Ring order (one camera, split launch): the same, with the lead set moved
// to the front, in centre-out order among themselves: the pixels within
// kRingWidth of the camera's critical curve (Bardeen's photon-orbit curve
// for the spin and inclination), whose rays orbit the hole, and the polar
// band, whose rays turn in theta within arcsin(kPolarSin) of the spin axis,
// where the step-size rule (integrator.c:109-122) shrinks the steps with
// min(X2, 1 - X2) (at 1024^2, a = 0.9375, i = 60: up to 1300 steps inside
// the grid against a mean of 200, in a column of pixels from the shadow to
// the image's edge, the last of which the centre-out order claims last),
// up to kLeadShare of the image in all (a camera near the axis would put
// nearly every pixel in the band).  The lead kernel traces them from the
// camera on a second stream from the start of the render, beside the
// approach kernel, instead of after it.
static constexpr double kRingWidth = 0.3;   // M: 12304 of 1024^2 pixels at 40 M, a = 0.9375, i = 60
static constexpr double kPolarSin = 0.02;   // 12022 more pixels there
static constexpr double kLeadShare = 0.05;  // of the image: the lead set's largest size

static const rt_ctx::OrderTable *find_order(rt_ctx *c, const long long key[kOrderKey])
{
    for (auto &t : c->orders) {
        if (memcmp(key, t.key, kOrderKey * sizeof(long long)) == 0) {
            t.used = ++c->use_clock;
            if (c->capturing)
                pin(c, t.d);
            return &t;
        }
    }
    return nullptr;
}

static void order_key(int W, int H, int T, double sx, double sy, bool ring, double a,
                      double incl, long long key[kOrderKey])
{
    const double v[4] = {sx, sy, ring ? a : 0., ring ? incl : 0.};
    key[0] = W;
    key[1] = H;
    key[2] = T | (ring ? 1 << 30 : 0);
    memcpy(key + 3, v, sizeof v);
}

// Pixels (row-major) within `width` (M) of the critical curve seen by a camera
// at inclination incl (degrees) around spin a: the impact parameters
// (alpha, beta) of the spherical photon orbits, alpha = -xi / sin i,
// beta^2 = eta + a^2 cos^2 i - xi^2 cot^2 i with xi(r), eta(r) of Bardeen
// (1973) for r between the prograde and retrograde circular photon orbits
// (a circle of radius sqrt(27) for a = 0), sampled densely and each sample
// marking the pixel centres within `width` of it.  Pixel centres as
// pixel_impact: alpha = -cam_x/2 + (x + 1/2) sx, beta = -cam_y/2 + (y + 1/2) sy.
// With polar > 0 also the pixels whose ray turns in theta closer to the axis
// than sin theta_min = polar: with the camera's constants of motion (E = 1,
// L = -alpha sin i, Carter Q = beta^2 + cos^2 i (alpha^2 - a^2)), the theta
// potential Q + a^2 cos^2 - L^2 cot^2 vanishes at u = cos^2 theta_min, the
// root in [0, 1] of a^2 u^2 + (Q + L^2 - a^2) u - Q = 0 (u = Q / (Q + L^2)
// for a = 0); rays with Q <= 0 never leave the equatorial band.  At most
// max_lead pixels in all: a camera near the axis sees most rays turn near it
// (all of them at i <= 1 deg), so beyond that count the band keeps its
// pixels of smallest sin theta_min (ties by pixel index).
static void ring_mask(int W, int H, double sx, double sy, double a, double incl,
                      double width, double polar, size_t max_lead, std::vector<char> &mask)
{
    mask.assign((size_t)W * H, 0);
    const double si = std::sin(incl * kPi / 180.), ci = std::cos(incl * kPi / 180.);
    auto mark = [&](double al, double be) {
        if (!(std::fabs(al) <= 1e9 && std::fabs(be) <= 1e9))   // off the image, or NaN
            return;
        const int x0 = std::max(0, (int)std::floor((al - width + 0.5 * W * sx) / sx - 0.5));
        const int x1 = std::min(W - 1, (int)std::ceil((al + width + 0.5 * W * sx) / sx - 0.5));
        const int y0 = std::max(0, (int)std::floor((be - width + 0.5 * H * sy) / sy - 0.5));
        const int y1 = std::min(H - 1, (int)std::ceil((be + width + 0.5 * H * sy) / sy - 0.5));
        for (int y = y0; y <= y1; y++) {
            const double dy = -0.5 * H * sy + (y + 0.5) * sy - be;
            for (int x = x0; x <= x1; x++) {
                const double dx = -0.5 * W * sx + (x + 0.5) * sx - al;
                if (dx * dx + dy * dy < width * width)
                    mask[(size_t)y * W + x] = 1;
            }
        }
    };
    const int n = 1 << 14;
    if (std::fabs(a) < 1e-4) {
        for (int q = 0; q < n; q++) {
            const double ph = 2. * kPi * q / n;
            mark(std::sqrt(27.) * std::cos(ph), std::sqrt(27.) * std::sin(ph));
        }
    } else {
        const double r1 = 2. * (1. + std::cos(2. / 3. * std::acos(-a)));
        const double r2 = 2. * (1. + std::cos(2. / 3. * std::acos(a)));
        const double lo = std::min(r1, r2), hi = std::max(r1, r2);
        // (alpha, beta^2) of the orbit at radius r
        auto curve = [&](double r, double &al, double &b2) {
            const double xi = (r * r * (3. - r) - a * a * (r + 1.)) / (a * (r - 1.));
            const double eta = r * r * r * (4. * a * a - r * (r - 3.) * (r - 3.)) /
                               (a * a * (r - 1.) * (r - 1.));
            al = -xi / si;
            b2 = eta + a * a * ci * ci - xi * xi * (ci * ci) / (si * si);
        };
        auto mark_r = [&](double r) {
            double al, b2;
            curve(r, al, b2);
            if (b2 >= 0.) {
                mark(al, std::sqrt(b2));
                mark(al, -std::sqrt(b2));
            }
        };
        // uniform in r, with samples inserted where consecutive points lie
        // more than width/4 apart (near the axis, i -> 0, the visible arc
        // maps from a narrow range of r) and the ends of the visible arc
        // (beta = 0) found by bisection
        double rp = lo, alp, b2p;
        curve(rp, alp, b2p);
        mark_r(rp);
        for (int q = 1; q <= n; q++) {
            const double r = lo + (hi - lo) * q / n;
            double al, b2;
            curve(r, al, b2);
            const bool vp = b2p >= 0., v = b2 >= 0.;
            // both points beyond the same edge of the image: nothing between
            // them to mark
            const double ex = 0.5 * W * sx + width, ey = 0.5 * H * sy + width;
            const bool beyond = (al > ex && alp > ex) || (al < -ex && alp < -ex) ||
                                (vp && v && b2 > ey * ey && b2p > ey * ey);
            if (vp && v && !beyond) {
                const double gap = std::hypot(al - alp, std::sqrt(b2) - std::sqrt(b2p));
                const double kk = std::ceil(gap / (0.25 * width));
                const int k = kk < 65536. ? (int)kk : 65536;   // NaN and inf too
                for (int t = 1; t < k; t++)
                    mark_r(rp + (r - rp) * t / k);
            } else if (vp != v) {
                double x0 = rp, x1 = r;   // b2 changes sign between them
                for (int it = 0; it < 60; it++) {
                    const double xm = 0.5 * (x0 + x1);
                    double alm, b2m;
                    curve(xm, alm, b2m);
                    if ((b2m >= 0.) == vp)
                        x0 = xm;
                    else
                        x1 = xm;
                }
                const double rb = vp ? x0 : x1, rv = vp ? rp : r;   // end, valid sample
                double alb, b2b, alv, b2v;
                curve(rb, alb, b2b);
                curve(rv, alv, b2v);
                mark(alb, 0.);
                // beta grows like sqrt(r - rb) from the end: samples at
                // r = rb + (rv - rb) s^2 are spaced evenly along the arc
                const double gap = std::hypot(alv - alb, std::sqrt(b2v));
                const double kk = std::ceil(gap / (0.25 * width));
                const int k = kk < 65536. ? (int)kk : 65536;
                for (int t = 1; t < k; t++) {
                    const double sq = (double)t / k;
                    mark_r(rb + (rv - rb) * sq * sq);
                }
            }
            mark_r(r);
            rp = r;
            alp = al;
            b2p = b2;
        }
    }
    if (!(polar > 0.))
        return;
    size_t nring = 0;
    for (char m : mask)
        nring += m != 0;
    std::vector<std::pair<double, int>> band;   // (sin^2 theta_min, pixel)
    for (int y = 0; y < H; y++) {
        const double be = -0.5 * H * sy + (y + 0.5) * sy;
        for (int x = 0; x < W; x++) {
            const double al = -0.5 * W * sx + (x + 0.5) * sx;
            const double L = -al * si, Q = be * be + ci * ci * (al * al - a * a);
            if (mask[(size_t)y * W + x] || !(Q > 0.))
                continue;
            double u;
            if (a * a < 1e-12) {
                u = Q / (Q + L * L);
            } else {
                // the root without cancellation: 2Q / (B + sqrt(B^2 + 4 a^2 Q))
                const double B = Q + L * L - a * a;
                const double sq = std::sqrt(B * B + 4. * a * a * Q);
                u = B > 0. ? 2. * Q / (B + sq) : (sq - B) / (2. * a * a);
            }
            if (1. - u < polar * polar)
                band.emplace_back(1. - u, y * W + x);
        }
    }
    const size_t room = max_lead > nring ? max_lead - nring : 0;
    if (band.size() > room) {
        std::nth_element(band.begin(), band.begin() + room, band.end());
        band.resize(room);
    }
    for (const auto &b : band)
        mask[(size_t)b.second] = 1;
}

// ring: incl and a give the camera; the table's first t->nlead entries are
// then the ring's pixels
static int ensure_order(rt_ctx *c, int W, int H, int T, double sx, double sy, bool ring,
                        double a, double incl, const rt_ctx::OrderTable **out)
{
    long long key[kOrderKey];
    order_key(W, H, T, sx, sy, ring, a, incl, key);
    if ((*out = find_order(c, key)))
        return RT_OK;
    if (c->capturing)
        return fail(c, RT_E_STATE, "the claim order for a new image shape is built by an eager "
                                   "render; render once before capturing");
    if (c->orders.size() >= kMaxOrderTables) {
        // free the least recently used table that no graph pins, once the
        // launches already queued on the stream (which may read it) are done
        size_t victim = c->orders.size();
        for (size_t i = 0; i < c->orders.size(); i++)
            if (c->orders[i].refs == 0 &&
                (victim == c->orders.size() || c->orders[i].used < c->orders[victim].used))
                victim = i;
        if (victim < c->orders.size()) {
            HIPCHK(c, hipStreamSynchronize(c->stream));
            HIPCHK(c, hipFree(c->orders[victim].d));
            c->orders.erase(c->orders.begin() + (long)victim);
        }
    }
    struct Item {
        double d;
        int tile, pix;
    };
    const size_t n = (size_t)W * H;
    std::vector<Item> items;
    std::vector<int> order;
    std::vector<char> lead;
    try {
        items.resize(n);
        order.resize(n);
        c->orders.reserve(c->orders.size() + 1);
        if (ring)
            ring_mask(W, H, sx, sy, a, incl, kRingWidth, kPolarSin,
                      (size_t)(kLeadShare * (double)n), lead);
    } catch (const std::bad_alloc &) {
        return fail(c, RT_E_NOMEM, "out of host memory for the claim order");
    }
    const int ntx = (W + T - 1) / T;
    for (int y = 0; y < H; y++) {
        const int ty = y / T;
        const double dy = (double)(ty * T + std::min(ty * T + T, H) - H) * sy;  // 2 (centre offset)
        for (int x = 0; x < W; x++) {
            const int tx = x / T;
            const double dx = (double)(tx * T + std::min(tx * T + T, W) - W) * sx;
            Item &e = items[(size_t)y * W + x];
            e.d = dx * dx + dy * dy;
            e.tile = ty * ntx + tx;
            e.pix = y * W + x;
        }
    }
    std::sort(items.begin(), items.end(), [](const Item &a, const Item &b) {
        if (a.d != b.d)
            return a.d < b.d;
        if (a.tile != b.tile)
            return a.tile < b.tile;
        return a.pix < b.pix;
    });
    size_t nlead = 0;
    if (ring) {
        for (size_t i = 0; i < n; i++)
            if (lead[items[i].pix])
                order[nlead++] = items[i].pix;
        size_t k = nlead;
        for (size_t i = 0; i < n; i++)
            if (!lead[items[i].pix])
                order[k++] = items[i].pix;
    } else {
        for (size_t i = 0; i < n; i++)
            order[i] = items[i].pix;
    }
    rt_ctx::OrderTable t;
    memcpy(t.key, key, sizeof key);
    t.d = nullptr;
    t.nlead = (int)nlead;
    t.refs = 0;
    t.used = ++c->use_clock;
    HIPCHK(c, hipMalloc((void **)&t.d, n * sizeof(int)));
    // on the context's stream, complete before `order` goes out of scope
    hipError_t e = hipMemcpyAsync(t.d, order.data(), n * sizeof(int), hipMemcpyHostToDevice,
                                  c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        (void)hipFree(t.d);
        return fail(c, RT_E_HIP, "claim order upload: %s", hipGetErrorString(e));
    }
    c->orders.push_back(t);
    *out = &c->orders.back();
    return RT_OK;
}

// Whether a render of ncam cameras and nf frequencies over npix pixels
// takes the ring order's lead launch: the ring order, one camera, one
// frequency (its re-trace at the gate runs inside the trace kernels), and
// the split launch of all rays in one chunk.
static bool ring_launch(const rt_ctx *c, long long npix, int ncam, int nf)
{
    const int pct = c->cfg.approach != 0 ? c->cfg.approach
                                         : (npix * ncam >= kApproachRays ? 100 : RT_APPROACH_NONE);
    return c->cfg.order == RT_ORDER_RING && ncam == 1 && nf == 1 && pct == 100 &&
           npix <= kMaxChunk;
}

// noalloc (rt_render_device_noalloc): the work queue, the FP64 re-trace's
// list and the spectrum's partial sums in the device's `scratch`; one fused
// launch of at most kNoAllocRays work items; a centre-out claim order only
// if one was built before for the image shape, tiled scan order otherwise.
static int render_impl(rt_ctx *c, const double *incl, int ncam, const double *freqs, int nf,
                       double *d_image, double *d_dimage, rt_ray_counters *d_counters,
                       double *d_spectrum, bool grad, bool noalloc = false)
{
    if (!c)
        return RT_E_INVALID;
    Params P;
    TraceArgs A;
    int rc = make_params(c, incl, ncam, freqs, nf, P);
    if (rc)
        return rc;
    if (!d_image || (grad && !d_dimage))
        return fail(c, RT_E_INVALID, "output buffers must not be NULL");
    if (grad && c->params.emission != RT_EMIS_THSYNCH)
        return fail(c, RT_E_INVALID, "gradients are available for the thermal synchrotron "
                                     "emission model (RT_EMIS_THSYNCH) only");
    if ((rc = select_device(c)))
        return rc;
    char *const scr = (char *)c->d_scratch;
    A.prims = c->d_prims;
    if (c->capturing)
        pin(c, c->d_prims);
    A.queue = noalloc ? (unsigned int *)(scr + kScratchQueue) : c->d_queue;
    A.refill = std::min(64, std::max(1, c->cfg.refill));
    A.tile = std::max(1, c->cfg.tile);
    A.order = nullptr;
    // the ring order's lead launch: one camera, one frequency (its re-trace
    // at the gate runs inside the trace kernels), the split launch in one
    // chunk, not for the no-allocation entry point
    const bool ring = !noalloc && ring_launch(c, P.npix, ncam, nf);
    int nlead = 0;
    if (c->cfg.order != RT_ORDER_SCAN) {
        const rt_ctx::OrderTable *t = nullptr;
        if (noalloc) {
            long long key[kOrderKey];
            order_key(P.width, P.height, A.tile, P.cam_x / P.width, P.cam_y / P.height, false, 0.,
                      0., key);
            t = find_order(c, key);
        } else {
            rc = ensure_order(c, P.width, P.height, A.tile, P.cam_x / P.width,
                              P.cam_y / P.height, ring, P.a, incl[0], &t);
            if (rc)
                return rc;
        }
        if (t) {
            A.order = t->d;
            nlead = ring ? t->nlead : 0;
        }
    }
    if (A.order) {
        A.ntx = 0;
        A.per_cam = P.npix;
    } else if (A.tile > 1) {
        A.ntx = (P.width + A.tile - 1) / A.tile;
        int nty = (P.height + A.tile - 1) / A.tile;
        long long pc = (long long)A.ntx * nty * A.tile * A.tile;
        if (pc * ncam > 0x7fffff00LL)
            return fail(c, RT_E_INVALID, "too many work items");
        A.per_cam = (int)pc;
    } else {
        A.ntx = 0;
        A.per_cam = P.npix;
    }
    A.unit = A.tile > 1 ? A.tile * A.tile : 64;
    const int total = ncam * A.per_cam;
    if (noalloc && total > kNoAllocRays)
        return fail(c, RT_E_INVALID, "rt_render_device_noalloc: %d work items (n_cam x W x H, "
                    "in whole claim tiles) exceed its scratch's %d", total, kNoAllocRays);
    A.image = d_image;
    A.dimage = grad ? d_dimage : nullptr;
    A.counters = d_counters;
    A.st = RayState{};
    if (noalloc) {
        A.xlist = (int2 *)(scr + kScratchXlist);
    } else {
        if (!c->d_xlist) {
            if (c->capturing)
                return fail(c, RT_E_STATE, "the FP64 re-trace's ray list is allocated by an "
                                           "eager render; render once before capturing");
            HIPCHK(c, hipMalloc((void **)&c->d_xlist, (size_t)kMaxChunk * sizeof(int2)));
        }
        A.xlist = c->d_xlist;
    }
    A.xcount = A.queue + 1;
    const int pct = noalloc ? RT_APPROACH_NONE
                    : c->cfg.approach != 0 ? c->cfg.approach
                                           : (total >= kApproachRays ? 100 : RT_APPROACH_NONE);
    // ring: the lead launch takes the table's first nlead pixels, the split
    // launch the rest
    TraceArgs L = A;
    if (nlead > 0) {
        L.order = A.order;
        L.per_cam = nlead;
        L.total = nlead;
        L.w0 = 0;
        L.nsaved = 0;
        L.st = RayState{};
        L.queue = c->d_queue + 2;
        L.xcount = L.queue + 1;
        A.order += nlead;
        A.per_cam = P.npix - nlead;
    }
    const int total_main = ncam * A.per_cam;
    if (pct > 0 &&
        (rc = ensure_state(c, (std::min<long long>(total_main, kMaxChunk) * pct + 99) / 100,
                           &A.st)))
        return rc;
    if (!c->capturing)
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    hipError_t e;
    bool d3 = c->grid.n3 > 1;
    const TraceArgs *lead = nlead > 0 ? &L : nullptr;
    if (grad)
        e = d3 ? dispatch_int<true, true>(c, P, A, total_main, pct, lead)
               : dispatch_int<true, false>(c, P, A, total_main, pct, lead);
    else
        e = d3 ? dispatch_int<false, true>(c, P, A, total_main, pct, lead)
               : dispatch_int<false, false>(c, P, A, total_main, pct, lead);
    HIPCHK(c, e);
    if (d_spectrum) {
        const int nb = (int)std::min<long long>(kSpecBlocks, ((long long)P.npix + 1023) / 1024);
        void (*part)(const double *, int, int, double *) =
            nf == 1 ? spectrum_partial<1>
                    : (nf <= 4 ? spectrum_partial<4>
                               : (nf <= 8 ? spectrum_partial<8> : spectrum_partial<16>));
        double *spec_part = noalloc ? (double *)(scr + kScratchSpec) : c->d_spec_part;
        hipLaunchKernelGGL(part, dim3((unsigned)nb, (unsigned)ncam), dim3(256), 0, c->stream,
                           (const double *)d_image, P.npix, nf, spec_part);
        hipLaunchKernelGGL(spectrum_final, dim3((unsigned)ncam), dim3(64), 0, c->stream,
                           (const double *)spec_part, nb, nf, d_spectrum);
        HIPCHK(c, hipGetLastError());
    }
    if (!c->capturing)
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    return RT_OK;
}

// Calls that allocate, copy or synchronise cannot be recorded into a graph.
static int no_capture(rt_ctx *c, const char *what)
{
    return c->capturing ? fail(c, RT_E_STATE, "%s cannot be called while capturing a graph", what)
                        : RT_OK;
}

static int render_host(rt_ctx *c, const double *incl, int ncam, const double *freqs, int nf,
                       double *h_image, double *h_dimage, double *h_spectrum, bool grad)
{
    if (!c)
        return RT_E_INVALID;
    if (int rc0 = no_capture(c, "rt_render"))
        return rc0;
    if (!h_image || (grad && !h_dimage))
        return fail(c, RT_E_INVALID, "output buffers must not be NULL");
    if (!c->have_params || !c->have_grid)
        return fail(c, RT_E_STATE, "rt_set_params and rt_load_prims must be called first");
    if (ncam < 1 || ncam > kMaxCam || nf < 1 || nf > kMaxFreq)
        return fail(c, RT_E_INVALID, "n_cam or n_freq out of range");
    int rc = select_device(c);
    if (rc)
        return rc;
    size_t n = (size_t)ncam * c->params.width * c->params.height * nf;
    if ((rc = ensure_buffer(c, &c->d_image, &c->image_cap, n * sizeof(double))))
        return rc;
    if (grad && (rc = ensure_buffer(c, &c->d_dimage, &c->dimage_cap, 2 * n * sizeof(double))))
        return rc;
    rc = render_impl(c, incl, ncam, freqs, nf, c->d_image, grad ? c->d_dimage : nullptr,
                     nullptr, h_spectrum ? c->d_spec : nullptr, grad);
    if (rc)
        return rc;
    HIPCHK(c, hipMemcpyAsync(h_image, c->d_image, n * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    if (grad)
        HIPCHK(c, hipMemcpyAsync(h_dimage, c->d_dimage, 2 * n * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
    if (h_spectrum)
        HIPCHK(c, hipMemcpyAsync(h_spectrum, c->d_spec, (size_t)ncam * nf * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    return RT_OK;
}

extern "C" {

const char *rt_version(void) { return "raptor_rt 0.1 (gfx950)"; }

rt_ctx *rt_create(int device)
{
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) {
        fail(nullptr, RT_E_HIP, "no HIP device available (%s)", hipGetErrorString(e));
        return nullptr;
    }
    if (device < 0 || device >= ndev) {
        fail(nullptr, RT_E_INVALID, "device %d out of range [0, %d)", device, ndev);
        return nullptr;
    }
    rt_ctx *c = new (std::nothrow) rt_ctx();
    if (!c) {
        fail(nullptr, RT_E_NOMEM, "out of host memory");
        return nullptr;
    }
    c->device = device;
    hipDeviceProp_t prop;
    if ((e = hipSetDevice(device)) != hipSuccess ||
        (e = hipGetDeviceProperties(&prop, device)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess ||
        (e = hipEventCreate(&c->evA)) != hipSuccess || (e = hipEventCreate(&c->evT)) != hipSuccess ||
        (e = hipEventCreate(&c->evX)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&c->lead_stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreate(&c->evL0)) != hipSuccess || (e = hipEventCreate(&c->evL1)) != hipSuccess ||
        (e = hipMalloc((void **)&c->d_queue, 4 * sizeof(unsigned int))) != hipSuccess) {
        fail(nullptr, RT_E_HIP, "HIP set-up failed: %s", hipGetErrorString(e));
        rt_destroy(c);
        return nullptr;
    }
    c->num_cu = prop.multiProcessorCount;
    if ((e = scratch_address(device, &c->d_scratch)) != hipSuccess) {
        fail(nullptr, RT_E_HIP, "the no-allocation scratch's address: %s", hipGetErrorString(e));
        rt_destroy(c);
        return nullptr;
    }
    if ((e = hipMalloc((void **)&c->d_spec_part,
                       (size_t)kMaxCam * kSpecBlocks * kMaxFreq * sizeof(double))) != hipSuccess ||
        (e = hipMalloc((void **)&c->d_spec, (size_t)kMaxCam * kMaxFreq * sizeof(double))) !=
            hipSuccess) {
        fail(nullptr, RT_E_HIP, "HIP set-up failed: %s", hipGetErrorString(e));
        rt_destroy(c);
        return nullptr;
    }
    std::vector<float2> tab;
    kappa_table(tab, c->ktab_u0, c->ktab_idu);
    c->ktab_n = (int)tab.size();
    if ((e = hipMalloc((void **)&c->d_ktab, tab.size() * sizeof(float2))) != hipSuccess ||
        (e = hipMemcpy(c->d_ktab, tab.data(), tab.size() * sizeof(float2),
                       hipMemcpyHostToDevice)) != hipSuccess) {
        fail(nullptr, RT_E_HIP, "HIP set-up failed: %s", hipGetErrorString(e));
        rt_destroy(c);
        return nullptr;
    }
    return c;
}


void rt_destroy(rt_ctx *c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->device);
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->d_prims);
    for (auto &r : c->retired_prims)
        (void)hipFree(r.d);
    (void)hipFree(c->d_queue);
    (void)hipFree(c->d_xlist);
    (void)hipFree(c->d_spec_part);
    (void)hipFree(c->d_spec);
    (void)hipFree(c->d_ktab);
    for (auto &t : c->orders)
        (void)hipFree(t.d);
    for (auto &b : c->states)
        (void)hipFree(b.d);
    (void)hipFree(c->d_image);
    (void)hipFree(c->d_dimage);
    if (c->ev0)
        (void)hipEventDestroy(c->ev0);
    if (c->ev1)
        (void)hipEventDestroy(c->ev1);
    for (hipEvent_t ev : {c->evA, c->evT, c->evX, c->ev_fork, c->ev_join, c->evL0, c->evL1})
        if (ev)
            (void)hipEventDestroy(ev);
    if (c->lead_stream) {
        (void)hipStreamSynchronize(c->lead_stream);
        (void)hipStreamDestroy(c->lead_stream);
    }
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *rt_last_error(const rt_ctx *c) { return c ? c->err : g_create_err; }

int rt_set_params(rt_ctx *c, const rt_params *p)
{
    if (!c)
        return RT_E_INVALID;
    if (!p)
        return fail(c, RT_E_INVALID, "params must not be NULL");
    if (p->width < 1 || p->height < 1 || (long long)p->width * p->height > 0x7fffff00LL)
        return fail(c, RT_E_INVALID, "bad image size %d x %d", p->width, p->height);
    if (!(p->MBH > 0.) || !(p->M_unit > 0.) || !(p->stepsize > 0.) || p->max_steps < 1 ||
        !(p->cam_size_x > 0.) || !(p->cam_size_y > 0.))
        return fail(c, RT_E_INVALID, "MBH, M_unit, stepsize, max_steps and camera size must be positive");
    if (p->integrator < RT_INT_RK2 || p->integrator > RT_INT_VERLET)
        return fail(c, RT_E_INVALID, "integrator must be RT_INT_RK2, RT_INT_RK4 or RT_INT_VERLET, "
                                     "got %d", p->integrator);
    if (p->emission < RT_EMIS_THSYNCH || p->emission > RT_EMIS_KAPPA)
        return fail(c, RT_E_INVALID, "emission must be RT_EMIS_THSYNCH, RT_EMIS_THSYNCHAV or "
                                     "RT_EMIS_KAPPA, got %d", p->emission);
    c->params = *p;
    c->have_params = true;
    return RT_OK;
}

int rt_load_prims(rt_ctx *c, const rt_grid *g, const float *h_prims)
{
    if (!c)
        return RT_E_INVALID;
    if (int rc0 = no_capture(c, "rt_load_prims"))
        return rc0;
    if (!g || !h_prims)
        return fail(c, RT_E_INVALID, "grid and prims must not be NULL");
    if (g->n1 < 2 || g->n2 < 2 || g->n3 < 1)
        return fail(c, RT_E_INVALID, "grid must be at least 2 x 2 x 1, got %d x %d x %d", g->n1,
                    g->n2, g->n3);
    if (g->R0 != 0.)
        return fail(c, RT_E_INVALID, "R0 != 0 is not supported (the MKS inverse metric of the "
                                     "reference assumes R0 = 0)");
    if (!(std::fabs(g->a) < 1.) || !(g->dx1 > 0.) || !(g->dx2 > 0.) ||
        (g->n3 > 1 && !(g->dx3 > 0.)))
        return fail(c, RT_E_INVALID, "need |a| < 1 and positive cell sizes");
    if ((long long)g->n1 * g->n2 * g->n3 > kMaxCells)
        return fail(c, RT_E_INVALID, "grids of more than 2^27 cells are not supported (%d x %d "
                    "x %d)", g->n1, g->n2, g->n3);
    int rc = select_device(c);
    if (rc)
        return rc;
    size_t ncell = (size_t)g->n1 * g->n2 * g->n3;
    if (c->d_prims) {
        if (c->prims_refs > 0) {
            try {
                c->retired_prims.push_back({c->d_prims, c->prims_refs});
            } catch (const std::bad_alloc &) {
                return fail(c, RT_E_NOMEM, "out of host memory");
            }
        } else {
            HIPCHK(c, hipStreamSynchronize(c->stream));   // renders still reading it
            HIPCHK(c, hipFree(c->d_prims));
        }
        c->d_prims = nullptr;
        c->prims_refs = 0;
    }
    c->have_grid = false;
    float *d_soa = nullptr;
    HIPCHK(c, hipMalloc((void **)&c->d_prims, ncell * 2 * sizeof(float4)));
    HIPCHK(c, hipMalloc((void **)&d_soa, ncell * 8 * sizeof(float)));
    hipError_t e = hipMemcpyAsync(d_soa, h_prims, ncell * 8 * sizeof(float),
                                  hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        unsigned blocks = (unsigned)((ncell + 255) / 256);
        hipLaunchKernelGGL(stage_prims, dim3(blocks), dim3(256), 0, c->stream, d_soa, ncell,
                           c->d_prims);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_soa);
    if (e != hipSuccess)
        return fail(c, RT_E_HIP, "primitive upload failed: %s", hipGetErrorString(e));
    c->grid = *g;
    c->have_grid = true;
    return RT_OK;
}

int rt_render(rt_ctx *c, const double *incl, int ncam, const double *freqs, int nf,
              double *h_image, double *h_spectrum)
{
    return render_host(c, incl, ncam, freqs, nf, h_image, nullptr, h_spectrum, false);
}

int rt_render_grad(rt_ctx *c, const double *incl, int ncam, const double *freqs, int nf,
                   double *h_image, double *h_dimage, double *h_spectrum)
{
    return render_host(c, incl, ncam, freqs, nf, h_image, h_dimage, h_spectrum, true);
}

int rt_render_device(rt_ctx *c, const double *incl, int ncam, const double *freqs, int nf,
                     double *d_image, double *d_dimage, rt_ray_counters *d_counters,
                     double *d_spectrum)
{
    if (!c)
        return RT_E_INVALID;
    return render_impl(c, incl, ncam, freqs, nf, d_image, d_dimage, d_counters, d_spectrum,
                       d_dimage != nullptr);
}

int rt_render_device_noalloc(rt_ctx *c, const double *incl, int ncam, const double *freqs,
                             int nf, double *d_image, double *d_dimage,
                             rt_ray_counters *d_counters, double *d_spectrum)
{
    if (!c)
        return RT_E_INVALID;
    if (!c->d_scratch)
        return fail(c, RT_E_STATE, "no scratch address for device %d", c->device);
    return render_impl(c, incl, ncam, freqs, nf, d_image, d_dimage, d_counters, d_spectrum,
                       d_dimage != nullptr, true);
}

int rt_synchronize(rt_ctx *c)
{
    if (!c)
        return RT_E_INVALID;
    if (int rc0 = no_capture(c, "rt_synchronize"))
        return rc0;
    int rc = select_device(c);
    if (rc)
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess)
        c->last_ms = ms;
    return RT_OK;
}

double rt_last_kernel_ms(const rt_ctx *c) { return c ? c->last_ms : 0.; }

int rt_last_kernel_times(rt_ctx *c, double *out)
{
    if (!c || !out)
        return RT_E_INVALID;
    if (int rc0 = no_capture(c, "rt_last_kernel_times"))
        return rc0;
    if (c->last_chunks != 1)
        return fail(c, RT_E_STATE, "the last render ran %d launch chunks; kernel times are "
                                   "kept for renders of one chunk", c->last_chunks);
    int rc = select_device(c);
    if (rc)
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const hipEvent_t ev[5] = {c->ev0, c->evA, c->evT, c->evX, c->ev1};
    for (int i = 0; i < 4; i++) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
        out[i] = ms;
    }
    out[4] = 0.;
    if (c->lead_timed) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->evL0, c->evL1));
        out[4] = ms;
    }
    return RT_OK;
}

double rt_jansky_factor(const rt_ctx *c)
{
    if (!c || !c->have_params)
        return 0.;
    const rt_params &p = c->params;
    double Rg = kGGrav * p.MBH / kC / kC;
    double dx = p.cam_size_x * Rg / kSourceDist;
    double dy = p.cam_size_y * Rg / kSourceDist;
    return 1.e23 * (dx / (double)p.width) * (dy / (double)p.height);
}

int rt_capture_begin(rt_ctx *c)
{
    if (!c)
        return RT_E_INVALID;
    if (c->capturing)
        return fail(c, RT_E_STATE, "a capture is already open");
    int rc = select_device(c);
    if (rc)
        return rc;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    c->capture_pins.clear();
    return RT_OK;
}

int rt_capture_end(rt_ctx *c, rt_graph **out)
{
    if (!c || !out)
        return RT_E_INVALID;
    if (!c->capturing)
        return fail(c, RT_E_STATE, "no capture is open");
    c->capturing = false;
    int rc = select_device(c);
    if (rc)
        return rc;
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamEndCapture(c->stream, &g));
    rt_graph *G = new (std::nothrow) rt_graph();
    if (!G) {
        (void)hipGraphDestroy(g);
        return fail(c, RT_E_NOMEM, "out of host memory");
    }
    G->graph = g;
    G->device = c->device;
    hipError_t e = hipGraphInstantiate(&G->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        delete G;
        return fail(c, RT_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    G->pins = c->capture_pins;
    auto pinned = [&](void *p) {
        return std::find(G->pins.begin(), G->pins.end(), p) != G->pins.end();
    };
    for (auto &t : c->orders)
        if (pinned(t.d))
            t.refs++;
    for (auto &b : c->states)
        if (pinned(b.d))
            b.refs++;
    if (c->d_prims && pinned(c->d_prims))
        c->prims_refs++;
    *out = G;
    return RT_OK;
}

int rt_graph_launch(rt_ctx *c, rt_graph *g)
{
    if (!c || !g)
        return RT_E_INVALID;
    if (g->device != c->device)
        return fail(c, RT_E_INVALID, "graph was captured on another device");
    if (int rc0 = no_capture(c, "rt_graph_launch"))
        return rc0;
    int rc = select_device(c);
    if (rc)
        return rc;
    c->last_chunks = 0;   // no per-kernel times for a replay
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, hipGraphLaunch(g->exec, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    return RT_OK;
}

int rt_graph_destroy(rt_ctx *c, rt_graph *g)
{
    if (!c)
        return RT_E_INVALID;
    if (!g)
        return RT_OK;
    int rc = select_device(c);
    if (rc)
        return rc;
    (void)hipStreamSynchronize(c->stream);
    if (g->exec)
        (void)hipGraphExecDestroy(g->exec);
    if (g->graph)
        (void)hipGraphDestroy(g->graph);
    auto pinned = [&](void *p) {
        return std::find(g->pins.begin(), g->pins.end(), p) != g->pins.end();
    };
    for (auto &t : c->orders)
        if (t.refs > 0 && pinned(t.d))
            t.refs--;
    for (auto &b : c->states)
        if (b.refs > 0 && pinned(b.d))
            b.refs--;
    if (c->prims_refs > 0 && c->d_prims && pinned(c->d_prims))
        c->prims_refs--;
    for (size_t i = 0; i < c->retired_prims.size();) {
        auto &r = c->retired_prims[i];
        if (r.refs > 0 && pinned(r.d))
            r.refs--;
        if (r.refs == 0) {
            (void)hipFree(r.d);
            c->retired_prims.erase(c->retired_prims.begin() + (long)i);
        } else {
            i++;
        }
    }
    // older state buffers that no graph uses any more (the stream is idle)
    for (size_t i = 0; i + 1 < c->states.size();) {
        if (c->states[i].refs == 0) {
            (void)hipFree(c->states[i].d);
            c->states.erase(c->states.begin() + (long)i);
        } else {
            i++;
        }
    }
    delete g;
    return RT_OK;
}

int rt_claim_order(rt_ctx *c, double inclination, int *h_order, int *n_lead)
{
    if (!c)
        return RT_E_INVALID;
    if (int rc0 = no_capture(c, "rt_claim_order"))
        return rc0;
    if (!h_order || !n_lead)
        return fail(c, RT_E_INVALID, "output buffers must not be NULL");
    Params P;
    const double freq = 2.3e11;
    int rc = make_params(c, &inclination, 1, &freq, 1, P);
    if (rc)
        return rc;
    if ((rc = select_device(c)))
        return rc;
    const size_t n = (size_t)P.npix;
    *n_lead = 0;
    if (c->cfg.order == RT_ORDER_SCAN) {
        // work_to_ray's scan: T x T tiles in row-major tile order, pixels
        // row-major inside a tile, items outside the image skipped
        const int T = std::max(1, c->cfg.tile), W = P.width, H = P.height;
        const int ntx = (W + T - 1) / T, nty = (H + T - 1) / T;
        size_t k = 0;
        for (int t = 0; t < ntx * nty; t++)
            for (int l = 0; l < T * T; l++) {
                const int x = (t % ntx) * T + l % T, y = (t / ntx) * T + l / T;
                if (x < W && y < H)
                    h_order[k++] = y * W + x;
            }
        return RT_OK;
    }
    const bool ring = ring_launch(c, P.npix, 1, 1);
    const rt_ctx::OrderTable *t = nullptr;
    if ((rc = ensure_order(c, P.width, P.height, std::max(1, c->cfg.tile), P.cam_x / P.width,
                           P.cam_y / P.height, ring, P.a, inclination, &t)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(h_order, t->d, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *n_lead = ring ? t->nlead : 0;
    return RT_OK;
}

int rt_set_launch_config(rt_ctx *c, const rt_launch_config *cfg)
{
    if (!c)
        return RT_E_INVALID;
    if (!cfg || (cfg->block != 0 && cfg->block != 64) || cfg->blocks_per_cu < 0 ||
        cfg->blocks_per_cu > 64 || cfg->refill < 0 || cfg->refill > 64 || cfg->tile < 0 ||
        cfg->tile > 64 ||
        (cfg->order != 0 && cfg->order != RT_ORDER_CENTRE && cfg->order != RT_ORDER_SCAN &&
         cfg->order != RT_ORDER_RING) ||
        cfg->approach < RT_APPROACH_NONE || cfg->approach > 100)
        return fail(c, RT_E_INVALID, "bad launch config (block 64, blocks_per_cu 0..64, "
                                     "refill 0..64, tile 0..64, order 0, RT_ORDER_CENTRE, "
                                     "RT_ORDER_SCAN or RT_ORDER_RING, approach RT_APPROACH_NONE "
                                     "or 0..100)");
    c->cfg.block = 64;
    c->cfg.blocks_per_cu = cfg->blocks_per_cu;
    c->cfg.refill = cfg->refill == 0 ? kDefaultRefill : cfg->refill;
    c->cfg.tile = cfg->tile == 0 ? kDefaultTile : cfg->tile;
    c->cfg.order = cfg->order == 0 ? RT_ORDER_RING : cfg->order;
    c->cfg.approach = cfg->approach;
    return RT_OK;
}

int rt_get_launch_config(const rt_ctx *c, rt_launch_config *cfg)
{
    if (!c || !cfg)
        return RT_E_INVALID;
    *cfg = c->cfg;
    return RT_OK;
}

int rt_device_alloc(rt_ctx *c, size_t bytes, void **d_ptr)
{
    if (!c || !d_ptr)
        return RT_E_INVALID;
    if (int rc0 = no_capture(c, "rt_device_alloc"))
        return rc0;
    int rc = select_device(c);
    if (rc)
        return rc;
    HIPCHK(c, hipMalloc(d_ptr, bytes));
    return RT_OK;
}

int rt_device_free(rt_ctx *c, void *d_ptr)
{
    if (!c)
        return RT_E_INVALID;
    if (int rc0 = no_capture(c, "rt_device_free"))
        return rc0;
    int rc = select_device(c);
    if (rc)
        return rc;
    HIPCHK(c, hipFree(d_ptr));
    return RT_OK;
}

int rt_memcpy_d2h(rt_ctx *c, void *h_dst, const void *d_src, size_t bytes)
{
    if (!c || (!h_dst && bytes) || (!d_src && bytes))
        return RT_E_INVALID;
    if (int rc0 = no_capture(c, "rt_memcpy_d2h"))
        return rc0;
    int rc = select_device(c);
    if (rc)
        return rc;
    HIPCHK(c, hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_fluid_params(rt_ctx *c, const double *h_X, int n, double *h_out)
{
    if (!c)
        return RT_E_INVALID;
    if (int rc0 = no_capture(c, "rt_fluid_params"))
        return rc0;
    if (n < 0 || (n && (!h_X || !h_out)))
        return fail(c, RT_E_INVALID, "bad point list");
    if (n == 0)
        return RT_OK;
    const double one = 2.3e11, incl = 60.;
    Params P;
    int rc = make_params(c, &incl, 1, &one, 1, P);
    if (rc)
        return rc;
    if ((rc = select_device(c)))
        return rc;
    double *d_in = nullptr, *d_out = nullptr;
    HIPCHK(c, hipMalloc((void **)&d_in, (size_t)n * 4 * sizeof(double)));
    hipError_t e = hipMalloc((void **)&d_out, (size_t)n * 14 * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_in, h_X, (size_t)n * 4 * sizeof(double), hipMemcpyHostToDevice,
                           c->stream);
    if (e == hipSuccess) {
        unsigned blocks = (unsigned)((n + 63) / 64);
        if (c->grid.n3 > 1)
            hipLaunchKernelGGL(probe_fluid<true>, dim3(blocks), dim3(64), 0, c->stream, P,
                               c->d_prims, d_in, n, d_out);
        else
            hipLaunchKernelGGL(probe_fluid<false>, dim3(blocks), dim3(64), 0, c->stream, P,
                               c->d_prims, d_in, n, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(h_out, d_out, (size_t)n * 14 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    if (e != hipSuccess)
        return fail(c, RT_E_HIP, "rt_fluid_params: %s", hipGetErrorString(e));
    return RT_OK;
}

int rt_probe(rt_ctx *c, int kind, const double *h_in, int n, double *h_out)
{
    if (!c)
        return RT_E_INVALID;
    if (int rc0 = no_capture(c, "rt_probe"))
        return rc0;
    if (n < 0 || (n && (!h_in || !h_out)))
        return fail(c, RT_E_INVALID, "bad point list");
    int nin, nout;
    switch (kind) {
    case RT_PROBE_METRIC: nin = 4; nout = 32; break;
    case RT_PROBE_ACCEL: nin = 8; nout = 4; break;
    case RT_PROBE_RK2: nin = 8; nout = 9; break;
    case RT_PROBE_CAMERA: nin = 3; nout = 8; break;
    case RT_PROBE_EMISSION: nin = 5; nout = 2; break;
    case RT_PROBE_EMISSION_MODEL: nin = 6; nout = 2; break;
    default: return fail(c, RT_E_INVALID, "unknown probe kind %d", kind);
    }
    if (n == 0)
        return RT_OK;
    Params P;
    memset(&P, 0, sizeof P);
    std::vector<double> in(h_in, h_in + (size_t)nin * n);
    if (kind == RT_PROBE_EMISSION_MODEL) {
        for (int i = 0; i < n; i++) {
            const double m = in[6 * i + 5];
            if (m != RT_EMIS_THSYNCH && m != RT_EMIS_THSYNCHAV && m != RT_EMIS_KAPPA)
                return fail(c, RT_E_INVALID, "unknown emission model %g", m);
        }
    } else if (kind != RT_PROBE_EMISSION) {
        if (!c->have_grid || !c->have_params)
            return fail(c, RT_E_STATE, "rt_set_params and rt_load_prims must be called first");
        std::vector<double> incl;
        if (kind == RT_PROBE_CAMERA) {
            // each point names its inclination; one camera per distinct value
            for (int i = 0; i < n; i++) {
                double v = in[3 * i + 2];
                auto it = std::find(incl.begin(), incl.end(), v);
                if (it == incl.end()) {
                    if ((int)incl.size() == kMaxCam)
                        return fail(c, RT_E_INVALID, "too many distinct inclinations");
                    incl.push_back(v);
                    it = incl.end() - 1;
                }
                in[3 * i + 2] = (double)(it - incl.begin());
            }
        } else {
            incl.push_back(60.);
        }
        const double nu = 2.3e11;
        int rc = make_params(c, incl.data(), (int)incl.size(), &nu, 1, P);
        if (rc)
            return rc;
    }
    int rc = select_device(c);
    if (rc)
        return rc;
    double *d_in = nullptr, *d_out = nullptr;
    HIPCHK(c, hipMalloc((void **)&d_in, in.size() * sizeof(double)));
    hipError_t e = hipMalloc((void **)&d_out, (size_t)n * nout * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_in, in.data(), in.size() * sizeof(double), hipMemcpyHostToDevice,
                           c->stream);
    if (e == hipSuccess) {
        dim3 grid((unsigned)((n + 63) / 64)), block(64);
        switch (kind) {
        case RT_PROBE_METRIC:
            hipLaunchKernelGGL(probe_metric, grid, block, 0, c->stream, P, d_in, n, d_out);
            break;
        case RT_PROBE_ACCEL:
            hipLaunchKernelGGL(probe_accel, grid, block, 0, c->stream, P, d_in, n, d_out);
            break;
        case RT_PROBE_RK2:
            hipLaunchKernelGGL(probe_rk2, grid, block, 0, c->stream, P, d_in, n, d_out);
            break;
        case RT_PROBE_CAMERA:
            hipLaunchKernelGGL(probe_camera, grid, block, 0, c->stream, P, d_in, n, d_out);
            break;
        case RT_PROBE_EMISSION_MODEL:
            hipLaunchKernelGGL(probe_emission_model, grid, block, 0, c->stream, kappa_with_table(c),
                               d_in, n, d_out);
            break;
        default:
            hipLaunchKernelGGL(probe_emission, grid, block, 0, c->stream, d_in, n, d_out);
            break;
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(h_out, d_out, (size_t)n * nout * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    if (e != hipSuccess)
        return fail(c, RT_E_HIP, "rt_probe: %s", hipGetErrorString(e));
    return RT_OK;
}

}  // extern "C"
