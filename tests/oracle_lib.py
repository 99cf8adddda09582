"""ctypes bindings of the CPU oracle (oracle/_build/libakaze_oracle.so).

Test infrastructure: imported by tests, by __graft_entry__.smoke() and by the
cpu_baseline leg of bench.py, never by the akaze_amd package.
"""
import ctypes
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "_build", "libakaze_oracle.so")
FEDREF_SO = os.path.join(REPO, "oracle", "_ref", "libfed_ref.so")

KP_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("size", "<f4"), ("angle", "<f4"),
                     ("response", "<f4"), ("octave", "<i4"), ("class_id", "<i4")])
DM_DTYPE = np.dtype([("queryIdx", "<i4"), ("trainIdx", "<i4"), ("imgIdx", "<i4"), ("distance", "<f4")])

c_float_p = ctypes.POINTER(ctypes.c_float)
c_int_p = ctypes.POINTER(ctypes.c_int)
c_u8_p = ctypes.POINTER(ctypes.c_uint8)


class Options(ctypes.Structure):
    _fields_ = [("omax", ctypes.c_int), ("nsublevels", ctypes.c_int), ("soffset", ctypes.c_float),
                ("derivative_factor", ctypes.c_float), ("diffusivity", ctypes.c_int),
                ("dthreshold", ctypes.c_float), ("min_dthreshold", ctypes.c_float),
                ("descriptor_pattern_size", ctypes.c_int), ("kcontrast_percentile", ctypes.c_float),
                ("kcontrast_nbins", ctypes.c_int), ("nms_max_waves", ctypes.c_int),
                ("descriptor", ctypes.c_int), ("descriptor_size", ctypes.c_int),
                ("descriptor_channels", ctypes.c_int)]


MAXL = 32


class Geom(ctypes.Structure):
    _fields_ = [("nlev", ctypes.c_int), ("omax", ctypes.c_int),
                ("w", ctypes.c_int * MAXL), ("h", ctypes.c_int * MAXL),
                ("octave", ctypes.c_int * MAXL), ("sublevel", ctypes.c_int * MAXL),
                ("esigma", ctypes.c_float * MAXL), ("etime", ctypes.c_float * MAXL),
                ("nsteps", ctypes.c_int * MAXL), ("tau", (ctypes.c_float * 256) * MAXL),
                ("deriv_step", ctypes.c_int * MAXL), ("border", ctypes.c_float * MAXL),
                ("det_size", ctypes.c_float * MAXL)]


class PropGeom(ctypes.Structure):
    _fields_ = [("nlev", ctypes.c_int), ("P", ctypes.c_int), ("tw", ctypes.c_int * MAXL), ("th", ctypes.c_int * MAXL),
                ("oct", ctypes.c_int * MAXL), ("base", ctypes.c_int * (MAXL + 1))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_options_default.argtypes = [ctypes.POINTER(Options)]
        L.orc_fed_tau_by_process_time.argtypes = [ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                                  c_float_p, ctypes.c_int]
        L.orc_geometry.argtypes = [ctypes.POINTER(Options), ctypes.c_int, ctypes.c_int, ctypes.POINTER(Geom)]
        L.orc_gauss_taps.argtypes = [ctypes.c_double, ctypes.c_int, c_float_p]
        L.orc_lowpass.argtypes = [c_float_p, c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.orc_contrast.restype = ctypes.c_float
        L.orc_contrast.argtypes = [c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                   c_int_p, c_float_p]
        L.orc_flow.argtypes = [c_float_p, c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float]
        L.orc_fed_step.argtypes = [c_float_p, c_float_p, c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_float]
        L.orc_halfsample.argtypes = [c_float_p, ctypes.c_int, ctypes.c_int, c_float_p, ctypes.c_int, ctypes.c_int]
        L.orc_hessian.argtypes = [c_float_p, c_float_p, c_float_p, c_float_p, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int]
        L.orc_find_extrema.argtypes = [c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_void_p, ctypes.c_int]
        L.orc_nms.argtypes = [ctypes.c_void_p, ctypes.c_int, c_int_p, ctypes.c_int, ctypes.c_int,
                              ctypes.c_void_p, c_int_p]
        L.orc_orientation.argtypes = [ctypes.c_void_p, c_float_p, c_float_p, ctypes.c_int, ctypes.c_int]
        L.orc_mldb.argtypes = [ctypes.c_void_p, c_float_p, c_float_p, c_float_p, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, c_u8_p, c_float_p]
        L.orc_match2nn.argtypes = [c_u8_p, ctypes.c_int, c_u8_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.orc_compare_tables.argtypes = [c_int_p, c_int_p]
        L.orc_mldb_ex.argtypes = [ctypes.c_void_p, c_float_p, c_float_p, c_float_p, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, c_int_p, c_int_p, ctypes.c_int, c_u8_p, c_float_p]
        L.orc_descriptor_subsample.restype = ctypes.c_int
        L.orc_descriptor_subsample.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, c_int_p, c_int_p]
        L.orc_mldb_table.restype = ctypes.c_int
        L.orc_mldb_table.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, c_int_p, c_int_p]
        L.orc_descriptor_bytes.restype = ctypes.c_int
        L.orc_descriptor_bytes.argtypes = [ctypes.c_void_p]
        L.orc_atan2f.restype = ctypes.c_float
        L.orc_atan2f.argtypes = [ctypes.c_float, ctypes.c_float]
        L.orc_sincosf.argtypes = [ctypes.c_float, c_float_p, c_float_p]
        L.orc_orient_weight.restype = ctypes.c_float
        L.orc_orient_weight.argtypes = [ctypes.c_int]
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(Options), ctypes.c_int, ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_get_geom.restype = ctypes.POINTER(Geom)
        L.orc_get_geom.argtypes = [ctypes.c_void_p]
        L.orc_scale_space.argtypes = [ctypes.c_void_p, c_float_p]
        L.orc_get_kcontrast.restype = ctypes.c_float
        L.orc_get_kcontrast.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_level.restype = c_float_p
        L.orc_level.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_detect.argtypes = [ctypes.c_void_p]
        L.orc_num_candidates.argtypes = [ctypes.c_void_p]
        L.orc_candidates.restype = ctypes.c_void_p
        L.orc_candidates.argtypes = [ctypes.c_void_p]
        L.orc_level_start.restype = c_int_p
        L.orc_level_start.argtypes = [ctypes.c_void_p]
        L.orc_nms_waves.argtypes = [ctypes.c_void_p]
        L.orc_num_keypoints.argtypes = [ctypes.c_void_p]
        L.orc_keypoints.restype = ctypes.c_void_p
        L.orc_keypoints.argtypes = [ctypes.c_void_p]
        L.orc_describe.argtypes = [ctypes.c_void_p]
        L.orc_descriptors.restype = ctypes.c_void_p
        L.orc_descriptors.argtypes = [ctypes.c_void_p]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_kept_candidates.restype = c_int_p
        L.orc_kept_candidates.argtypes = [ctypes.c_void_p]
        L.orc_prop_make_geom.argtypes = [ctypes.c_void_p, ctypes.POINTER(PropGeom)]
        L.orc_prop_features.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p]
        L.orc_prop_labels.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.orc_prop_loss_grad.restype = ctypes.c_double
        L.orc_prop_loss_grad.argtypes = [ctypes.POINTER(PropGeom), ctypes.c_int, c_float_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_float, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_void_p]
        L.orc_prop_adam.argtypes = [c_float_p, c_float_p, c_float_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_int]
        _lib = L
    return _lib


def fptr(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(c_float_p)


def default_options(**kw):
    o = Options()
    lib().orc_options_default(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def geometry(W, H, **kw):
    g = Geom()
    n = lib().orc_geometry(ctypes.byref(default_options(**kw)), W, H, ctypes.byref(g))
    assert n > 0
    return g


def fed_tau(T, M=1, tau_max=0.25, reordering=True):
    buf = (ctypes.c_float * 256)()
    n = lib().orc_fed_tau_by_process_time(T, M, tau_max, int(reordering), buf, 256)
    return np.array(buf[:n], dtype=np.float32)


def lowpass(img, var, ksize):
    out = np.empty_like(img)
    lib().orc_lowpass(fptr(img), fptr(out), img.shape[1], img.shape[0], var, ksize)
    return out


def gauss_taps(var, R):
    buf = (ctypes.c_float * 11)()
    lib().orc_gauss_taps(var, R, buf)
    return np.array(buf[:2 * R + 1], dtype=np.float32)


def contrast(img, perc=0.7, nbins=300):
    hist = np.zeros(512, np.int32)
    mx = ctypes.c_float()
    k = lib().orc_contrast(fptr(img), img.shape[1], img.shape[0], perc, nbins,
                           hist.ctypes.data_as(c_int_p), ctypes.byref(mx))
    return np.float32(k), hist[:min(nbins, 512)], np.float32(mx.value)


def flow(L, kcontrast, diffusivity=1):
    out = np.empty_like(L)
    lib().orc_flow(fptr(L), fptr(out), L.shape[1], L.shape[0], diffusivity, kcontrast)
    return out


def fed_step(L, g, tau):
    out = np.empty_like(L)
    lib().orc_fed_step(fptr(L), fptr(g), fptr(out), L.shape[1], L.shape[0], tau)
    return out


def halfsample(img, ow, oh):
    out = np.empty((oh, ow), np.float32)
    lib().orc_halfsample(fptr(img), img.shape[1], img.shape[0], fptr(out), ow, oh)
    return out


def hessian(Lsm, step):
    h, w = Lsm.shape
    lx = np.empty_like(Lsm); ly = np.empty_like(Lsm); det = np.empty_like(Lsm)
    lib().orc_hessian(fptr(Lsm), fptr(lx), fptr(ly), fptr(det), w, h, step)
    return lx, ly, det


def find_extrema(det, border, thr, level, octave, size, cap=100000):
    out = np.zeros(cap, KP_DTYPE)
    n = lib().orc_find_extrema(fptr(det), det.shape[1], det.shape[0], border, thr, level, octave, size,
                               out.ctypes.data, cap)
    return out[:min(n, cap)]


def nms(cand, level_start, max_waves=10000):
    cand = np.ascontiguousarray(cand, KP_DTYPE)
    ls = np.ascontiguousarray(level_start, np.int32)
    out = np.zeros(max(len(cand), 1), KP_DTYPE)
    waves = ctypes.c_int()
    n = lib().orc_nms(cand.ctypes.data, len(cand), ls.ctypes.data_as(c_int_p), len(ls) - 1, max_waves,
                      out.ctypes.data, ctypes.byref(waves))
    return out[:n], waves.value


def orient_describe(pipe, kps, patsize=10):
    """Orientation and MLDB (oracle) of keypoints on a Pipeline's levels."""
    kps = np.ascontiguousarray(kps, KP_DTYPE).copy()
    desc = np.zeros((len(kps), 61), np.uint8)
    L = lib()
    g = pipe.geom
    for i in range(len(kps)):
        l = int(kps[i]["class_id"])
        w, h = g.w[l], g.h[l]
        kp = ctypes.c_void_p(kps.ctypes.data + 28 * i)
        lx, ly, lt = L.orc_level(pipe.ctx, l, 2), L.orc_level(pipe.ctx, l, 3), L.orc_level(pipe.ctx, l, 0)
        L.orc_orientation(kp, lx, ly, w, h)
        L.orc_mldb(kp, lt, lx, ly, w, h, patsize, desc[i].ctypes.data_as(c_u8_p), None)
    return kps, desc


def match2nn(q, t):
    q = np.ascontiguousarray(q, np.uint8)
    t = np.ascontiguousarray(t, np.uint8)
    assert q.shape[1] == t.shape[1]
    out = np.zeros(2 * len(q), DM_DTYPE)
    lib().orc_match2nn(q.ctypes.data_as(c_u8_p), len(q), t.ctypes.data_as(c_u8_p), len(t), q.shape[1],
                       out.ctypes.data)
    return out.reshape(-1, 2)


def compare_tables():
    a = np.zeros(488, np.int32); b = np.zeros(488, np.int32)
    lib().orc_compare_tables(a.ctypes.data_as(c_int_p), b.ctypes.data_as(c_int_p))
    return a, b


def descriptor_subsample(nbits, patsize=10, nch=3):
    """orc_descriptor_subsample: (samples n x 3 [grid, x, y], comparisons nbits x 2)."""
    smp = np.zeros(3 * 29, np.int32)
    cmp = np.zeros(2 * 486, np.int32)
    n = lib().orc_descriptor_subsample(nbits, patsize, nch, smp.ctypes.data_as(c_int_p), cmp.ctypes.data_as(c_int_p))
    if n < 0:
        return None
    return smp[:3 * n].reshape(n, 3), cmp[:2 * nbits].reshape(nbits, 2)


def mldb_table(descriptor_size=0, patsize=10, nch=3):
    """orc_mldb_table: (idx1, idx2) into the 87 cell values, one pair per bit."""
    a = np.zeros(488, np.int32)
    b = np.zeros(488, np.int32)
    n = lib().orc_mldb_table(descriptor_size, patsize, nch, a.ctypes.data_as(c_int_p), b.ctypes.data_as(c_int_p))
    if n < 0:
        return None
    return a[:n], b[:n]


def mldb_ex(kp, Lt, Lx, Ly, patsize, nch, table):
    """orc_mldb_ex on one keypoint: (descriptor bytes, 87 values)."""
    a, b = (np.ascontiguousarray(t, np.int32) for t in table)
    h, w = Lt.shape
    desc = np.zeros((len(a) + 7) // 8, np.uint8)
    vals = np.zeros(87, np.float32)
    lib().orc_mldb_ex(np.ascontiguousarray(kp, KP_DTYPE).ctypes.data, fptr(Lt), fptr(Lx), fptr(Ly), w, h, patsize,
                      nch, a.ctypes.data_as(c_int_p), b.ctypes.data_as(c_int_p), len(a),
                      desc.ctypes.data_as(c_u8_p), fptr(vals))
    return desc, vals


def atan2f(y, x):
    return np.float32(lib().orc_atan2f(y, x))


def sincosf(x):
    s = ctypes.c_float(); c = ctypes.c_float()
    lib().orc_sincosf(x, ctypes.byref(s), ctypes.byref(c))
    return np.float32(s.value), np.float32(c.value)


class Pipeline:
    """The whole CPU oracle pipeline for one image size (AKAZE.cpp:211-449)."""

    def __init__(self, W, H, threads=0, **kw):
        L = lib()
        if threads:
            L.orc_set_threads(threads)
        self.opts = default_options(**kw)
        self.ctx = L.orc_create(ctypes.byref(self.opts), W, H)
        if not self.ctx:
            raise ValueError("bad geometry")
        self.geom = L.orc_get_geom(self.ctx).contents
        self.W, self.H = W, H

    def __del__(self):
        if getattr(self, "ctx", None) and _lib is not None:
            _lib.orc_destroy(self.ctx)
            self.ctx = None

    @property
    def nlev(self):
        return self.geom.nlev

    def scale_space(self, img):
        img = np.ascontiguousarray(img, np.float32)
        assert img.shape == (self.H, self.W)
        lib().orc_scale_space(self.ctx, fptr(img))

    def level(self, lvl, which):
        w, h = self.geom.w[lvl], self.geom.h[lvl]
        p = lib().orc_level(self.ctx, lvl, which)
        return np.ctypeslib.as_array(p, shape=(h, w)).copy()

    def kcontrast(self, lvl):
        return np.float32(lib().orc_get_kcontrast(self.ctx, lvl))

    def detect(self):
        return lib().orc_detect(self.ctx)

    def candidates(self):
        L = lib()
        n = L.orc_num_candidates(self.ctx)
        if n == 0:
            return np.zeros(0, KP_DTYPE)
        buf = (ctypes.c_char * (n * 28)).from_address(L.orc_candidates(self.ctx))
        return np.frombuffer(bytes(buf), KP_DTYPE).copy()

    def level_start(self):
        p = lib().orc_level_start(self.ctx)
        return np.array([p[i] for i in range(self.nlev + 1)], np.int32)

    def waves(self):
        return lib().orc_nms_waves(self.ctx)

    def keypoints(self):
        L = lib()
        n = L.orc_num_keypoints(self.ctx)
        if n == 0:
            return np.zeros(0, KP_DTYPE)
        buf = (ctypes.c_char * (n * 28)).from_address(L.orc_keypoints(self.ctx))
        return np.frombuffer(bytes(buf), KP_DTYPE).copy()

    def describe(self):
        L = lib()
        n = L.orc_describe(self.ctx)
        nb = L.orc_descriptor_bytes(self.ctx)
        if n == 0:
            return np.zeros((0, nb), np.uint8)
        buf = (ctypes.c_char * (n * nb)).from_address(L.orc_descriptors(self.ctx))
        return np.frombuffer(bytes(buf), np.uint8).reshape(n, nb).copy()

    def prop_geom(self):
        g = PropGeom()
        lib().orc_prop_make_geom(self.ctx, ctypes.byref(g))
        return g

    def prop_features(self):
        g = self.prop_geom()
        out = np.zeros(g.P * 4, np.uint16)
        lib().orc_prop_features(self.ctx, ctypes.c_float(max(self.opts.dthreshold, self.opts.min_dthreshold)),
                                out.ctypes.data)
        return out

    def prop_labels(self):
        g = self.prop_geom()
        out = np.zeros(g.P, np.uint8)
        lib().orc_prop_labels(self.ctx, out.ctypes.data)
        return out

    def run(self, img):
        self.scale_space(img)
        self.detect()
        desc = self.describe()
        return self.keypoints(), desc


PROP_NPARAM = 14513
# flat parameter layout of the proposal network (offsets into the vector)
PROP_OFF = {"W1": 0, "b1": 576, "W2": 592, "b2": 5200, "W3": 5232, "b3": 14448, "w4": 14480, "b4": 14512}


def make_prop_geom(levels):
    """PropGeom from [(tw, th, octave), ...] tile maps."""
    g = PropGeom()
    g.nlev = len(levels)
    P = 0
    for i, (tw, th, o) in enumerate(levels):
        g.tw[i], g.th[i], g.oct[i], g.base[i] = tw, th, o, P
        P += tw * th
    g.base[len(levels)] = P
    g.P = P
    return g


def prop_loss_grad(g, nimg, params, x0, labels, wpos=1.0, emulate=True, want_grad=True):
    """(loss, grad float64 or None, prob float32) of the proposal oracle."""
    params = np.ascontiguousarray(params, np.float32)
    x0 = np.ascontiguousarray(x0, np.uint16)
    lab = None if labels is None else np.ascontiguousarray(labels, np.uint8)
    grad = np.zeros(PROP_NPARAM, np.float64) if want_grad else None
    prob = np.zeros(g.P * nimg, np.float32)
    loss = lib().orc_prop_loss_grad(ctypes.byref(g), nimg, fptr(params), x0.ctypes.data,
                                    lab.ctypes.data if lab is not None else None, ctypes.c_float(wpos),
                                    1 if emulate else 0, grad.ctypes.data if grad is not None else None,
                                    prob.ctypes.data)
    return loss, grad, prob


def prop_adam(params, m, v, grad, lr, step):
    p = np.ascontiguousarray(params, np.float32).copy()
    m = np.ascontiguousarray(m, np.float32).copy()
    v = np.ascontiguousarray(v, np.float32).copy()
    g = np.ascontiguousarray(grad, np.float64)
    lib().orc_prop_adam(fptr(p), fptr(m), fptr(v), g.ctypes.data, len(p), ctypes.c_float(lr), step)
    return p, m, v


def f32_to_bf16(a):
    """Round-to-nearest-even bfloat16 bits of float32 values."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_to_f32(b):
    return (np.ascontiguousarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)
