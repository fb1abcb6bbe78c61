"""Python host binding of the C-ABI in include/raptor_rt.h (ctypes).

This mirrors the reference's calculate_image (core.c:74-152) and its model
hooks: :class:`Tracer` holds one rt_ctx (one GPU), ``load_dump`` plays
init_model's storage role (raptor_harm_model.c:59-66) and ``render`` is the
pixel loop.  There is no CPU path: if libraptor_rt.so is missing or no GPU is
visible, the constructor raises.
"""
from __future__ import annotations

import ctypes
import os
import numpy as np

from .torus import Dump

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libraptor_rt.so")

RT_OK = 0
RT_MAX_FREQ = 16
RT_MAX_CAMERAS = 64
PROBE_METRIC, PROBE_ACCEL, PROBE_RK2, PROBE_CAMERA, PROBE_EMISSION = 1, 2, 3, 4, 5
PROBE_EMISSION_MODEL = 6

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


class RtGrid(ctypes.Structure):
    _fields_ = [("n1", ctypes.c_int), ("n2", ctypes.c_int), ("n3", ctypes.c_int),
                ("startx1", ctypes.c_double), ("startx2", ctypes.c_double),
                ("startx3", ctypes.c_double), ("dx1", ctypes.c_double),
                ("dx2", ctypes.c_double), ("dx3", ctypes.c_double), ("a", ctypes.c_double),
                ("hslope", ctypes.c_double), ("R0", ctypes.c_double)]


class RtParams(ctypes.Structure):
    _fields_ = [("MBH", ctypes.c_double), ("M_unit", ctypes.c_double),
                ("R_high", ctypes.c_double), ("R_low", ctypes.c_double),
                ("width", ctypes.c_int), ("height", ctypes.c_int),
                ("cam_size_x", ctypes.c_double), ("cam_size_y", ctypes.c_double),
                ("stepsize", ctypes.c_double), ("max_steps", ctypes.c_int),
                ("integrator", ctypes.c_int), ("emission", ctypes.c_int)]


class RtLaunchConfig(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int), ("blocks_per_cu", ctypes.c_int),
                ("refill", ctypes.c_int), ("tile", ctypes.c_int), ("order", ctypes.c_int),
                ("approach", ctypes.c_int)]


# rt_params.integrator (include/raptor_rt.h): int_method of parameters.h:186-189
INT_RK2, INT_RK4, INT_VERLET = 0, 1, 2
# rt_params.emission: the thermal fit of the path, the angle-averaged thermal
# fit and the kappa fits of j_nu.c
EMIS_THSYNCH, EMIS_THSYNCHAV, EMIS_KAPPA = 0, 1, 2

# claim orders of rt_launch_config.order (include/raptor_rt.h)
ORDER_CENTRE = 1
ORDER_SCAN = 2
ORDER_RING = 3

# rt_launch_config.approach: percent of the rays whose camera-to-grid leg runs
# in a kernel of its own (1..100), or none
APPROACH_NONE = -1
NOALLOC_MAX_RAYS = 1 << 22   # RT_NOALLOC_MAX_RAYS


class RtError(RuntimeError):
    pass


# (name, restype, argtypes) of every function the header declares
SIGNATURES = [
    ("rt_create", _vp, [ctypes.c_int]),
    ("rt_destroy", None, [_vp]),
    ("rt_last_error", ctypes.c_char_p, [_vp]),
    ("rt_version", ctypes.c_char_p, []),
    ("rt_set_params", ctypes.c_int, [_vp, ctypes.POINTER(RtParams)]),
    ("rt_load_prims", ctypes.c_int, [_vp, ctypes.POINTER(RtGrid), ctypes.POINTER(ctypes.c_float)]),
    ("rt_render", ctypes.c_int, [_vp, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]),
    ("rt_render_grad", ctypes.c_int, [_vp, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp]),
    ("rt_render_device", ctypes.c_int, [_vp, _dp, ctypes.c_int, _dp, ctypes.c_int, _vp, _vp, _vp,
                                         _vp]),
    ("rt_render_device_noalloc", ctypes.c_int, [_vp, _dp, ctypes.c_int, _dp, ctypes.c_int, _vp,
                                                 _vp, _vp, _vp]),
    ("rt_synchronize", ctypes.c_int, [_vp]),
    ("rt_last_kernel_ms", ctypes.c_double, [_vp]),
    ("rt_last_kernel_times", ctypes.c_int, [_vp, _dp]),
    ("rt_jansky_factor", ctypes.c_double, [_vp]),
    ("rt_set_launch_config", ctypes.c_int, [_vp, ctypes.POINTER(RtLaunchConfig)]),
    ("rt_claim_order", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_void_p,
                                      ctypes.POINTER(ctypes.c_int)]),
    ("rt_get_launch_config", ctypes.c_int, [_vp, ctypes.POINTER(RtLaunchConfig)]),
    ("rt_capture_begin", ctypes.c_int, [_vp]),
    ("rt_capture_end", ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    ("rt_graph_launch", ctypes.c_int, [_vp, _vp]),
    ("rt_graph_destroy", ctypes.c_int, [_vp, _vp]),
    ("rt_device_alloc", ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    ("rt_device_free", ctypes.c_int, [_vp, _vp]),
    ("rt_memcpy_d2h", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    ("rt_fluid_params", ctypes.c_int, [_vp, _dp, ctypes.c_int, _dp]),
    ("rt_probe", ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int, _dp]),
]

_LIB = None


def load_library(path=LIB_PATH):
    """Load libraptor_rt.so and declare its signatures (raises if missing)."""
    global _LIB
    if _LIB is not None and path == LIB_PATH:
        return _LIB
    if not os.path.exists(path):
        raise RtError("libraptor_rt.so not built (%s); run `python -m raptor_amd.build`" % path)
    lib = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _LIB = lib
    return lib


def _d(a):
    return a.ctypes.data_as(_dp)


class Graph:
    """A captured sequence of render_device calls (rt_capture_begin/end)."""

    def __init__(self, tracer, handle):
        self.tracer, self.handle = tracer, handle

    def launch(self):
        self.tracer._check(self.tracer.lib.rt_graph_launch(self.tracer.ctx, self.handle))

    def destroy(self):
        if self.handle:
            self.tracer._check(self.tracer.lib.rt_graph_destroy(self.tracer.ctx, self.handle))
            self.handle = None


class Tracer:
    """One GPU context.  Typical use::

        tr = Tracer()
        tr.set_params(width=1024, height=1024, M_unit=1e23, R_high=10.)
        tr.load_dump(dump)
        image, spectrum = tr.render([60.], [2.3e11])
    """

    def __init__(self, device=0, lib_path=LIB_PATH):
        self.lib = load_library(lib_path)
        self.ctx = self.lib.rt_create(device)
        if not self.ctx:
            raise RtError("rt_create failed: %s" % self.lib.rt_last_error(None).decode())
        self.width = self.height = None

    def close(self):
        if self.ctx:
            self.lib.rt_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != RT_OK:
            raise RtError("%d: %s" % (rc, self.lib.rt_last_error(self.ctx).decode()))

    # -- set-up ---------------------------------------------------------------
    def set_params(self, *, width, height, M_unit, R_high, R_low=1.0, MBH=8.9505e39,
                   cam_size=40.0, cam_size_y=None, stepsize=0.03, max_steps=100000,
                   integrator=INT_RK2, emission=EMIS_THSYNCH):
        p = RtParams(MBH=MBH, M_unit=M_unit, R_high=R_high, R_low=R_low, width=width,
                     height=height, cam_size_x=cam_size,
                     cam_size_y=cam_size if cam_size_y is None else cam_size_y,
                     stepsize=stepsize, max_steps=max_steps, integrator=integrator,
                     emission=emission)
        self._check(self.lib.rt_set_params(self.ctx, ctypes.byref(p)))
        self.width, self.height = width, height
        self.params = p

    def load_dump(self, dump: Dump):
        g = dump.grid
        grid = RtGrid(n1=g.n1, n2=g.n2, n3=g.n3, startx1=g.startx1, startx2=g.startx2,
                      startx3=g.startx3, dx1=g.dx1, dx2=g.dx2, dx3=g.dx3, a=g.a,
                      hslope=g.hslope, R0=g.R0)
        prims = np.ascontiguousarray(dump.prims, dtype=np.float32)
        self._check(self.lib.rt_load_prims(self.ctx, ctypes.byref(grid),
                                           prims.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def set_launch_config(self, blocks_per_cu=0, refill=0, tile=0, order=0, approach=0):
        """order: ORDER_RING, ORDER_CENTRE, ORDER_SCAN or 0 (library default, ring);
        approach: percent of the rays (in claim order) whose camera-to-grid leg
        runs in a kernel of its own, 1..100, APPROACH_NONE or 0 (library
        default)."""
        cfg = RtLaunchConfig(block=64, blocks_per_cu=blocks_per_cu, refill=refill, tile=tile,
                             order=order, approach=approach)
        self._check(self.lib.rt_set_launch_config(self.ctx, ctypes.byref(cfg)))

    def claim_order(self, inclination):
        """(order, n_lead): the pixels in the order a one-camera, one-frequency
        render at this inclination claims them, the first n_lead traced by the
        ring order's lead kernel."""
        order = np.zeros(self.width * self.height, np.int32)
        nl = ctypes.c_int(0)
        self._check(self.lib.rt_claim_order(self.ctx, float(inclination), order.ctypes.data,
                                            ctypes.byref(nl)))
        return order, nl.value

    def get_launch_config(self):
        cfg = RtLaunchConfig()
        self._check(self.lib.rt_get_launch_config(self.ctx, ctypes.byref(cfg)))
        return dict(block=cfg.block, blocks_per_cu=cfg.blocks_per_cu, refill=cfg.refill,
                    tile=cfg.tile, order=cfg.order, approach=cfg.approach)

    # -- the pixel loop -----------------------------------------------------------
    def render(self, inclinations, freqs, grad=False):
        """Returns image [ncam, H*W, nf] and spectrum [ncam, nf]; with grad also
        dimage [ncam, H*W, nf, 2] (d/d ln M_unit, d/d R_high)."""
        incl = np.ascontiguousarray(np.atleast_1d(inclinations), float)
        fr = np.ascontiguousarray(np.atleast_1d(freqs), float)
        npix = self.width * self.height
        img = np.empty((incl.size, npix, fr.size))
        spec = np.empty((incl.size, fr.size))
        if grad:
            dimg = np.empty((incl.size, npix, fr.size, 2))
            self._check(self.lib.rt_render_grad(self.ctx, _d(incl), incl.size, _d(fr), fr.size,
                                                _d(img), _d(dimg), _d(spec)))
            return img, dimg, spec
        self._check(self.lib.rt_render(self.ctx, _d(incl), incl.size, _d(fr), fr.size,
                                       _d(img), _d(spec)))
        return img, spec

    def render_device(self, inclinations, freqs, d_image, d_dimage=None, d_counters=None,
                      d_spectrum=None):
        """The pixel loop into device buffers (addresses from alloc): image
        [ncam, H*W, nf], optional dimage [ncam, H*W, nf, 2], per-ray counters
        [ncam*H*W] and spectrum [ncam, nf]; returns without synchronising."""
        incl = np.ascontiguousarray(np.atleast_1d(inclinations), float)
        fr = np.ascontiguousarray(np.atleast_1d(freqs), float)
        self._check(self.lib.rt_render_device(self.ctx, _d(incl), incl.size, _d(fr), fr.size,
                                              d_image, d_dimage, d_counters, d_spectrum))

    def render_device_noalloc(self, inclinations, freqs, d_image, d_dimage=None,
                              d_counters=None, d_spectrum=None):
        """render_device without any device allocation (the library's static
        scratch, at most NOALLOC_MAX_RAYS work items); the same outputs."""
        incl = np.ascontiguousarray(np.atleast_1d(inclinations), float)
        fr = np.ascontiguousarray(np.atleast_1d(freqs), float)
        self._check(self.lib.rt_render_device_noalloc(self.ctx, _d(incl), incl.size, _d(fr),
                                                      fr.size, d_image, d_dimage, d_counters,
                                                      d_spectrum))

    def synchronize(self):
        self._check(self.lib.rt_synchronize(self.ctx))

    def last_kernel_ms(self):
        return self.lib.rt_last_kernel_ms(self.ctx)

    def last_kernel_times(self):
        """ms of the last render's approach, trace, FP64 re-trace and spectrum
        kernels and of the ring order's lead kernel, which runs beside them
        (renders of one launch chunk)."""
        out = np.zeros(5)
        self._check(self.lib.rt_last_kernel_times(self.ctx, _d(out)))
        return dict(zip(("approach", "trace", "exact", "spectrum", "lead"), out))

    def jansky_factor(self):
        return self.lib.rt_jansky_factor(self.ctx)

    # -- graphs ------------------------------------------------------------------
    def capture_begin(self):
        self._check(self.lib.rt_capture_begin(self.ctx))

    def capture_end(self):
        g = _vp()
        self._check(self.lib.rt_capture_end(self.ctx, ctypes.byref(g)))
        return Graph(self, g.value)

    # -- device buffers -----------------------------------------------------------
    def alloc(self, nbytes):
        p = _vp()
        self._check(self.lib.rt_device_alloc(self.ctx, nbytes, ctypes.byref(p)))
        return p.value

    def free(self, ptr):
        self._check(self.lib.rt_device_free(self.ctx, ptr))

    def d2h(self, ptr, shape, dtype):
        out = np.empty(shape, dtype)
        self._check(self.lib.rt_memcpy_d2h(self.ctx, out.ctypes.data, ptr, out.nbytes))
        return out

    # -- device model hook and probes -------------------------------------------
    def fluid_params(self, X):
        X = np.ascontiguousarray(np.atleast_2d(X), float)
        out = np.empty((X.shape[0], 14))
        self._check(self.lib.rt_fluid_params(self.ctx, _d(X), X.shape[0], _d(out)))
        return out

    def probe(self, kind, inp):
        nin = {PROBE_METRIC: 4, PROBE_ACCEL: 8, PROBE_RK2: 8, PROBE_CAMERA: 3,
               PROBE_EMISSION: 5, PROBE_EMISSION_MODEL: 6}[kind]
        nout = {PROBE_METRIC: 32, PROBE_ACCEL: 4, PROBE_RK2: 9, PROBE_CAMERA: 8,
                PROBE_EMISSION: 2, PROBE_EMISSION_MODEL: 2}[kind]
        inp = np.ascontiguousarray(np.asarray(inp, float).reshape(-1, nin))
        out = np.empty((inp.shape[0], nout))
        self._check(self.lib.rt_probe(self.ctx, kind, _d(inp), inp.shape[0], _d(out)))
        return out
