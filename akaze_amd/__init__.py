"""MI355X A-KAZE: Python bindings of libakaze_amd.so (ctypes over the C-ABI).

Carries the reference's Python module (python/python_bindings.cpp:27-40):
AKAZEOptions with setWidth/setHeight, AKAZE with Create_Nonlinear_Scale_Space,
Feature_Detection (n x 7 float32 keypoints) and Compute_Descriptors (desc
n x 61 uint8, kpts n x 7 float32 in [x, y, size, angle, response, octave,
class_id] order, AKAZE.cpp:290-298), and Matcher.BFMatch returning n x 8
float32; akaze_amd.libakaze_pybindings exposes exactly those names.  The
structured-array forms are AKAZE.detect / describe / run and
Matcher.match2nn.  The evaluation helpers of src/lib/utils.cpp (PGM input,
save_keypoints, read_homography, NNDR, homography and RANSAC inliers) are
module functions.

The HIP library must be built (python -m akaze_amd.build); there is no CPU
fallback: every call runs the HIP kernels or raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libakaze_amd.so")

KP_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("size", "<f4"), ("angle", "<f4"),
                     ("response", "<f4"), ("octave", "<i4"), ("class_id", "<i4")])
DM_DTYPE = np.dtype([("queryIdx", "<i4"), ("trainIdx", "<i4"), ("imgIdx", "<i4"), ("distance", "<f4")])

STATUS = {0: "ok", 1: "invalid argument", 2: "HIP runtime error", 3: "out of memory",
          4: "capacity exceeded, output truncated", 5: "call order violated", 6: "option not implemented",
          7: "no HIP device"}
AKZ_ERR_OVERFLOW = 4


class AkazeError(RuntimeError):
    def __init__(self, status, where):
        super().__init__("%s: %s (status %d)" % (where, STATUS.get(status, "?"), status))
        self.status = status


class COptions(ctypes.Structure):
    _fields_ = [("omax", ctypes.c_int), ("nsublevels", ctypes.c_int), ("img_width", ctypes.c_int),
                ("img_height", ctypes.c_int), ("soffset", ctypes.c_float), ("derivative_factor", ctypes.c_float),
                ("sderivatives", ctypes.c_float), ("diffusivity", ctypes.c_int), ("dthreshold", ctypes.c_float),
                ("min_dthreshold", ctypes.c_float), ("descriptor", ctypes.c_int), ("descriptor_size", ctypes.c_int),
                ("descriptor_channels", ctypes.c_int), ("descriptor_pattern_size", ctypes.c_int),
                ("kcontrast", ctypes.c_float), ("kcontrast_percentile", ctypes.c_float),
                ("kcontrast_nbins", ctypes.c_int), ("maxkeypoints", ctypes.c_int), ("nms_max_waves", ctypes.c_int),
                ("verbosity", ctypes.c_int), ("fed_steps_per_launch", ctypes.c_int)]


_OPTION_FIELDS = frozenset(f[0] for f in COptions._fields_)


class CTiming(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("kcontrast", "scale", "derivatives", "detector", "extrema", "subpixel", "descriptor")]


_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_SZ = ctypes.c_size_t
_U64 = ctypes.c_uint64
_D = ctypes.c_double

# every exported C-ABI symbol with its argument types (include/akaze_mi.h)
SIGNATURES = {
    "akz_options_default": [_P],
    "akz_status_string": [_I],
    "akz_device_count": [_P],
    "akz_create": [_P, _P],
    "akz_destroy": [_P],
    "akz_num_levels": [_P, _P],
    "akz_descriptor_bytes": [_P, _P, _P],
    "akz_level_size": [_P, _I, _P, _P, _P, _P, _P],
    "akz_build_scale_space": [_P, _P, _SZ],
    "akz_build_scale_space_device": [_P, _P, _SZ],
    "akz_detect": [_P, _P],
    "akz_describe": [_P, _P, _P, _I, _P],
    "akz_process_device_async": [_P, _P, _SZ],
    "akz_wait": [_P],
    "akz_results_device": [_P, _P, _P, _P],
    "akz_read_results": [_P, _P, _P, _I, _P],
    "akz_input_buffer": [_P, _P, _P],
    "akz_set_graph_mode": [_P, _I],
    "akz_set_keep_determinant": [_P, _I],
    "akz_set_suppression_pool": [_P, _I],
    "akz_read_level": [_P, _I, _I, _P],
    "akz_read_kcontrast": [_P, _I, _P],
    "akz_read_candidates": [_P, _P, _I, _P, _P, _P],
    "akz_get_timing": [_P, _P],
    "akz_profile_fed": [_P, _P, _SZ, _P, _P, _P, _P],
    "akz_copy_kernel_run": [_P, _P, _SZ, _I, _P],
    "akz_create_batch": [_P, _I, _P],
    "akz_batch_capacity": [_P, _P],
    "akz_batch_input_buffer": [_P, _P, _P, _P],
    "akz_process_batch_device_async": [_P, _P, _SZ, _SZ, _I],
    "akz_read_batch_counts": [_P, _I, _P, _P],
    "akz_read_batch_results": [_P, _I, _P, _P, _I, _P],
    "akz_batch_results_device": [_P, _I, _P, _P, _P],
    "akz_read_batch_level": [_P, _I, _I, _I, _P],
    "akz_profile_fed_batch": [_P, _P, _SZ, _SZ, _I, _P, _P, _P, _P],
    "akz_prop_num_params": [_P],
    "akz_prop_create": [_U64, _P],
    "akz_prop_destroy": [_P],
    "akz_prop_set_params": [_P, _P],
    "akz_prop_get_params": [_P, _P],
    "akz_prop_positions": [_P, _P],
    "akz_prop_extract": [_P, _I, _P, _P],
    "akz_prop_forward": [_P, _P, _P, _P, _I, _P],
    "akz_prop_train_step": [_P, _P, _P, _P, _P, _I, _F, _F, _P, _P],
    "akz_set_proposal": [_P, _P, _F],
    "akz_prop_read_probs": [_P, _I, _P],
    "akz_synth_device_batch": [_U64, _I, _I, _I, _P, _SZ, _SZ],
    "akz_synth_device_warp": [_P, _I, _I, _SZ, _P, _F, _U64, _P, _SZ],
    "akz_autotune": [_P, _I, _I],
    "akz_get_fed_variants": [_P, _P],
    "akz_set_fed_variants": [_P, _P],
    "akz_ragged_create": [_P, _I, _P, _P, _P, _P],
    "akz_ragged_destroy": [_P],
    "akz_ragged_context": [_P, _I, _P],
    "akz_ragged_set_graph_mode": [_P, _I],
    "akz_ragged_autotune": [_P, _I],
    "akz_ragged_process": [_P, _I, _P, _P, _P, _P, _P],
    "akz_ragged_set_work_lists": [_P, _I],
    "akz_ragged_read": [_P, _I, _P, _P, _I, _P],
    "akz_ragged_results_device": [_P, _I, _P, _P, _P],
    "akz_matcher_create": [_P],
    "akz_matcher_destroy": [_P],
    "akz_match2nn": [_P, _P, _I, _P, _I, _SZ, _P],
    "akz_match2nn_device": [_P, _P, _I, _P, _I, _P],
    "akz_match2nn_device_noalloc": [_P, _I, _P, _I, _P, _P],
    "akz_matcher_wait": [_P],
    "akz_matcher_last_ms": [_P, _P],
    "akz_dev_alloc": [_SZ, _P],
    "akz_dev_free": [_P],
    "akz_memcpy_h2d": [_P, _P, _SZ],
    "akz_memcpy_d2h": [_P, _P, _SZ],
    "akz_memcpy_d2d": [_P, _P, _SZ],
    "akz_memcpy_2d": [_P, _SZ, _P, _SZ, _SZ, _SZ],
    "akz_device_synchronize": [],
    "akz_selftest_reciprocal": [_P],
    "akz_synth_image": [_U64, _I, _I, _P],
    "akz_synth_homography": [_U64, _I, _I, _D, _D, _D, _D, _P],
    "akz_synth_warp": [_P, _I, _I, _P, _D, _U64, _P],
    "akz_read_pgm": [ctypes.c_char_p, _P, _SZ, _P, _P],
    "akz_save_keypoints": [ctypes.c_char_p, _P, _I, _P, _I, _I],
    "akz_read_homography": [ctypes.c_char_p, _P],
    "akz_matches2points_nndr": [_P, _I, _P, _I, _P, _I, ctypes.c_float, _P, _P],
    "akz_inliers_homography": [_P, _I, _P, ctypes.c_float, _P, _P],
    "akz_inliers_ransac": [_P, _I, ctypes.c_float, _U64, _I, _D, _P, _P, _P],
}

_lib = None


def lib():
    """The loaded HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libakaze_amd.so is not built: run `python -m akaze_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_char_p if name == "akz_status_string" else ctypes.c_int
        _lib = L
    return _lib


def _check(st, where, allow=()):
    if st != 0 and st not in allow:
        raise AkazeError(st, where)
    return st


def device_count():
    n = ctypes.c_int(0)
    lib().akz_device_count(ctypes.byref(n))
    return n.value


# ---------------------------------------------------------------- synthetic inputs
def synth_image(seed, w, h):
    """Seeded synthetic uint8 image (blobs, shapes, 1/f noise)."""
    out = np.empty((h, w), np.uint8)
    _check(lib().akz_synth_image(seed, w, h, out.ctypes.data), "akz_synth_image")
    return out


def synth_homography(seed, w, h, max_rot_deg=30.0, min_scale=0.7, max_scale=1.3, max_persp=1e-4):
    H = np.empty(9, np.float64)
    _check(lib().akz_synth_homography(seed, w, h, max_rot_deg, min_scale, max_scale, max_persp, H.ctypes.data),
           "akz_synth_homography")
    return H.reshape(3, 3)


def synth_warp(img, H, noise_std=0.0, noise_seed=0):
    img = np.ascontiguousarray(img, np.uint8)
    Hc = np.ascontiguousarray(H, np.float64)
    out = np.empty_like(img)
    _check(lib().akz_synth_warp(img.ctypes.data, img.shape[1], img.shape[0], Hc.ctypes.data, noise_std, noise_seed,
                                out.ctypes.data), "akz_synth_warp")
    return out


def to_float(img_u8):
    """uint8 -> float32 in [0, 1] (the reference's input normalisation)."""
    return img_u8.astype(np.float32) / np.float32(255.0)


# ---------------------------------------------------------------- I/O and evaluation (src/lib/utils.cpp)
def _path(p):
    return os.fsencode(os.fspath(p))


def read_pgm(path):
    """cv::imread(path, 0) for PGM (P5/P2) files: uint8 H x W."""
    w, h = ctypes.c_int(), ctypes.c_int()
    _check(lib().akz_read_pgm(_path(path), None, 0, ctypes.byref(w), ctypes.byref(h)), "akz_read_pgm")
    out = np.empty((h.value, w.value), np.uint8)
    _check(lib().akz_read_pgm(_path(path), out.ctypes.data, out.nbytes, ctypes.byref(w), ctypes.byref(h)),
           "akz_read_pgm")
    return out


def save_keypoints(path, kpts, desc=None, save_desc=True):
    """save_keypoints (utils.cpp:97-151).  kpts: KP_DTYPE array or n x 7
    float32 [x, y, size, angle, response, octave, class_id]."""
    kps = _as_keypoints(kpts)
    d = np.zeros((len(kps), 0), np.uint8) if desc is None else np.ascontiguousarray(desc, np.uint8)
    _check(lib().akz_save_keypoints(_path(path), kps.ctypes.data, len(kps), d.ctypes.data, d.shape[1] if d.ndim == 2
                                    else 0, int(bool(save_desc))), "akz_save_keypoints")


def read_homography(path):
    """read_homography (utils.cpp:340-380): 3 x 3 float64 (float precision), or None."""
    H = np.zeros(9, np.float64)
    st = lib().akz_read_homography(_path(path), H.ctypes.data)
    return None if st != 0 else H.reshape(3, 3)


def matches2points_nndr(train, query, matches, nndr=0.8):
    """matches2points_nndr (utils.cpp:154-170): k x 4 float32 rows
    (train[queryIdx].pt, query[trainIdx].pt) of the rows with best < nndr * second."""
    tr, qu = _as_keypoints(train), _as_keypoints(query)
    m = np.ascontiguousarray(matches, DM_DTYPE).reshape(-1, 2)
    out = np.zeros((max(len(m), 1), 4), np.float32)
    k = ctypes.c_int()
    _check(lib().akz_matches2points_nndr(tr.ctypes.data, len(tr), qu.ctypes.data, len(qu), m.ctypes.data, len(m),
                                         nndr, out.ctypes.data, ctypes.byref(k)), "akz_matches2points_nndr")
    return out[:k.value].copy()


def inliers_homography(pairs, H, min_error=2.5):
    """compute_inliers_homography (utils.cpp:203-243): boolean mask."""
    p = np.ascontiguousarray(pairs, np.float32).reshape(-1, 4)
    Hd = np.ascontiguousarray(H, np.float64).reshape(9)
    mask = np.zeros(max(len(p), 1), np.uint8)
    k = ctypes.c_int()
    _check(lib().akz_inliers_homography(p.ctypes.data, len(p), Hd.ctypes.data, min_error, mask.ctypes.data,
                                        ctypes.byref(k)), "akz_inliers_homography")
    return mask[:len(p)].astype(bool)


def inliers_ransac(pairs, error=2.5, seed=0x5EED, max_iters=2000, confidence=0.995):
    """compute_inliers_ransac (utils.cpp:173-200) with the library's own
    RANSAC homography: (H 3 x 3, boolean mask)."""
    p = np.ascontiguousarray(pairs, np.float32).reshape(-1, 4)
    H = np.zeros(9, np.float64)
    mask = np.zeros(max(len(p), 1), np.uint8)
    k = ctypes.c_int()
    _check(lib().akz_inliers_ransac(p.ctypes.data, len(p), error, seed, max_iters, confidence, H.ctypes.data,
                                    mask.ctypes.data, ctypes.byref(k)), "akz_inliers_ransac")
    return H.reshape(3, 3), mask[:len(p)].astype(bool)


def kpts_to_mat(kps):
    """KeyPoint array -> n x 7 float32 [x, y, size, angle, response, octave,
    class_id], the order of mat2kpvec (AKAZE.cpp:290-298)."""
    kps = np.asarray(kps, KP_DTYPE)
    return np.stack([kps["x"], kps["y"], kps["size"], kps["angle"], kps["response"],
                     kps["octave"].astype(np.float32), kps["class_id"].astype(np.float32)],
                    axis=1).astype(np.float32).reshape(-1, 7)


def mat_to_kpts(mat):
    """n x 7 float32 -> KeyPoint array (mat2kpvec, AKAZE.cpp:290-298)."""
    mat = np.asarray(mat, np.float32).reshape(-1, 7)
    kps = np.zeros(len(mat), KP_DTYPE)
    for i, n in enumerate(("x", "y", "size", "angle", "response")):
        kps[n] = mat[:, i]
    kps["octave"] = mat[:, 5].astype(np.int32)
    kps["class_id"] = mat[:, 6].astype(np.int32)
    return kps


def _as_keypoints(k):
    a = np.asarray(k)
    if a.dtype == KP_DTYPE:
        return np.ascontiguousarray(a)
    return mat_to_kpts(a)


# ---------------------------------------------------------------- device memory
class DeviceBuffer:
    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        _check(lib().akz_dev_alloc(nbytes, ctypes.byref(p)), "akz_dev_alloc")
        self.ptr = p.value
        self.nbytes = nbytes

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _check(lib().akz_memcpy_h2d(self.ptr, arr.ctypes.data, arr.nbytes), "akz_memcpy_h2d")

    def download(self, arr):
        assert arr.flags.c_contiguous and arr.nbytes <= self.nbytes
        _check(lib().akz_memcpy_d2h(arr.ctypes.data, self.ptr, arr.nbytes), "akz_memcpy_d2h")
        return arr

    def free(self):
        if self.ptr and _lib is not None:
            _lib.akz_dev_free(self.ptr)
        self.ptr = None

    def __del__(self):
        self.free()


# ---------------------------------------------------------------- API mirror
class AKAZEOptions:
    """AKAZEOptions (src/lib/AKAZEConfig.h:84-143) with the reference defaults."""

    def __init__(self, **kw):
        c = COptions()
        lib().akz_options_default(ctypes.byref(c))
        self._c = c
        for k, v in kw.items():
            setattr(self, k, v)

    def __getattr__(self, k):
        if k.startswith("_"):
            raise AttributeError(k)
        return getattr(self._c, k)

    def __setattr__(self, k, v):
        if k == "_c":
            object.__setattr__(self, k, v)
        elif k in _OPTION_FIELDS:
            setattr(self._c, k, v)
        else:
            raise AttributeError("unknown AKAZE option %r" % k)

    def setWidth(self, w):
        self._c.img_width = w

    def setHeight(self, h):
        self._c.img_height = h


class AKAZE:
    """libAKAZECU::AKAZE (src/lib/AKAZE.h:58-143) over the HIP library."""

    def __init__(self, options):
        self.options = options
        p = ctypes.c_void_p()
        _check(lib().akz_create(ctypes.byref(options._c), ctypes.byref(p)), "akz_create")
        self.ctx = p.value
        n = ctypes.c_int()
        lib().akz_num_levels(self.ctx, ctypes.byref(n))
        self.nlev = n.value
        self.W, self.H = options.img_width, options.img_height
        self.nump = 0
        nb = ctypes.c_int()
        _check(lib().akz_descriptor_bytes(self.ctx, ctypes.byref(nb), None), "akz_descriptor_bytes")
        self.desc_bytes = nb.value

    def close(self):
        if getattr(self, "ctx", None) and _lib is not None:
            _lib.akz_destroy(self.ctx)
        self.ctx = None

    def __del__(self):
        self.close()

    def level_size(self, lvl):
        w, h, o, s = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        e = ctypes.c_float()
        _check(lib().akz_level_size(self.ctx, lvl, ctypes.byref(w), ctypes.byref(h), ctypes.byref(o),
                                    ctypes.byref(s), ctypes.byref(e)), "akz_level_size")
        return w.value, h.value, o.value, s.value, e.value

    def set_suppression_pool(self, entries):
        """Cap the suppression's LDS pool for long neighbour lists (0..4096);
        images whose lists do not fit take the global-memory path."""
        _check(lib().akz_set_suppression_pool(self.ctx, int(entries)), "akz_set_suppression_pool")

    # --- reference API
    def Create_Nonlinear_Scale_Space(self, img):
        img = np.ascontiguousarray(img, np.float32)
        if img.shape != (self.H, self.W):
            return -1
        _check(lib().akz_build_scale_space(self.ctx, img.ctypes.data, img.strides[0]), "akz_build_scale_space")
        self._detected = False
        return 0

    def detect(self):
        """AKAZE::Feature_Detection on the device; returns the kept count."""
        n = ctypes.c_int()
        st = lib().akz_detect(self.ctx, ctypes.byref(n))
        _check(st, "akz_detect", allow=(AKZ_ERR_OVERFLOW,))
        self.nump = n.value
        self.overflow = st == AKZ_ERR_OVERFLOW
        self._detected = True
        return self.nump

    def describe(self):
        """AKAZE::Compute_Descriptors; returns (kpts KP_DTYPE array, desc
        n x desc_bytes uint8: 61 for the default 486-bit M-LDB)."""
        cap = max(self.nump, 1)
        kps = np.zeros(cap, KP_DTYPE)
        desc = np.zeros((cap, self.desc_bytes), np.uint8)
        n = ctypes.c_int()
        st = lib().akz_describe(self.ctx, kps.ctypes.data, desc.ctypes.data, cap, ctypes.byref(n))
        _check(st, "akz_describe", allow=(AKZ_ERR_OVERFLOW,))
        m = min(n.value, cap)
        return kps[:m].copy(), desc[:m].copy()

    def detected_keypoints(self):
        """The detector's kept keypoints (angle 0 until described)."""
        cap = max(self.nump, 1)
        kps = np.zeros(cap, KP_DTYPE)
        n = ctypes.c_int()
        st = lib().akz_read_results(self.ctx, kps.ctypes.data, None, cap, ctypes.byref(n))
        _check(st, "akz_read_results", allow=(AKZ_ERR_OVERFLOW,))
        return kps[:min(n.value, cap)].copy()

    def Feature_Detection(self):
        """Python-binding form (python_bindings.cpp:34, AKAZE.cpp:302-310):
        n x 7 float32 keypoints [x, y, size, angle, response, octave,
        class_id].  The reference returns an unfilled matrix (its kpvec2mat
        is empty, AKAZE.cpp:279-287); this one holds the kept keypoints."""
        self.detect()
        return kpts_to_mat(self.detected_keypoints())

    def Compute_Descriptors(self):
        """Python-binding form (python_bindings.cpp:35, AKAZE.cpp:381-395):
        (desc n x 61 uint8, kpts n x 7 float32).  Detection runs first unless
        it already ran on the current scale space (the reference runs it
        again)."""
        if not getattr(self, "_detected", False):
            self.detect()
        kps, desc = self.describe()
        return desc, kpts_to_mat(kps)

    Feature_Detection_ = Feature_Detection
    Compute_Descriptors_ = Compute_Descriptors

    def Get_Computation_Times(self):
        t = CTiming()
        _check(lib().akz_get_timing(self.ctx, ctypes.byref(t)), "akz_get_timing")
        return {n: getattr(t, n) for n, _ in CTiming._fields_}

    # --- debug / parity readback
    def read_level(self, lvl, which):
        w, h = self.level_size(lvl)[:2]
        out = np.empty((h, w), np.float32)
        _check(lib().akz_read_level(self.ctx, lvl, which, out.ctypes.data), "akz_read_level")
        return out

    def read_kcontrast(self, lvl):
        k = ctypes.c_float()
        _check(lib().akz_read_kcontrast(self.ctx, lvl, ctypes.byref(k)), "akz_read_kcontrast")
        return np.float32(k.value)

    def read_candidates(self, cap=1 << 20):
        n = ctypes.c_int()
        w = ctypes.c_int()
        ls = np.zeros(self.nlev + 1, np.int32)
        st = lib().akz_read_candidates(self.ctx, None, 0, ctypes.byref(n), ls.ctypes.data, ctypes.byref(w))
        _check(st, "akz_read_candidates", allow=(AKZ_ERR_OVERFLOW,))
        out = np.zeros(max(n.value, 1), KP_DTYPE)
        _check(lib().akz_read_candidates(self.ctx, out.ctypes.data, n.value, ctypes.byref(n), ls.ctypes.data,
                                         ctypes.byref(w)), "akz_read_candidates")
        return out[:n.value], ls, w.value

    # --- throughput path
    def input_buffer(self):
        p = ctypes.c_void_p()
        pitch = ctypes.c_size_t()
        _check(lib().akz_input_buffer(self.ctx, ctypes.byref(p), ctypes.byref(pitch)), "akz_input_buffer")
        return p.value, pitch.value

    def set_graph_mode(self, on):
        _check(lib().akz_set_graph_mode(self.ctx, int(bool(on))), "akz_set_graph_mode")

    def process_device_async(self, d_img, pitch_bytes):
        _check(lib().akz_process_device_async(self.ctx, d_img, pitch_bytes), "akz_process_device_async")

    def wait(self):
        _check(lib().akz_wait(self.ctx), "akz_wait")

    def profile_fed(self, d_img, pitch_bytes):
        """(FED ms per image, algorithmic bytes, launches, must-move bytes)
        from HIP events on the context stream."""
        ms = ctypes.c_float(); by = ctypes.c_double(); n = ctypes.c_int(); mb = ctypes.c_double()
        _check(lib().akz_profile_fed(self.ctx, d_img, pitch_bytes, ctypes.byref(ms), ctypes.byref(by),
                                     ctypes.byref(n), ctypes.byref(mb)), "akz_profile_fed")
        return ms.value, by.value, n.value, mb.value

    def read_results(self, cap):
        kps = np.zeros(max(cap, 1), KP_DTYPE)
        desc = np.zeros((max(cap, 1), self.desc_bytes), np.uint8)
        n = ctypes.c_int()
        st = lib().akz_read_results(self.ctx, kps.ctypes.data, desc.ctypes.data, cap, ctypes.byref(n))
        _check(st, "akz_read_results", allow=(AKZ_ERR_OVERFLOW,))
        m = min(n.value, cap)
        return kps[:m].copy(), desc[:m].copy(), n.value

    def run(self, img):
        """Scale space + detection + description of one host image."""
        if self.Create_Nonlinear_Scale_Space(img) != 0:
            raise ValueError("image shape does not match the options")
        self.detect()
        return self.describe()


class BatchAKAZE:
    """Same-size image batches on one context: every stage is one launch over
    all images of a call.  No counterpart in the reference, which processes
    one image per AKAZE object; results equal AKAZE's per image."""

    def __init__(self, options, batch):
        self.options = options
        p = ctypes.c_void_p()
        _check(lib().akz_create_batch(ctypes.byref(options._c), int(batch), ctypes.byref(p)), "akz_create_batch")
        self.ctx = p.value
        self.batch = int(batch)
        self.W, self.H = options.img_width, options.img_height
        n = ctypes.c_int()
        lib().akz_num_levels(self.ctx, ctypes.byref(n))
        self.nlev = n.value
        _check(lib().akz_descriptor_bytes(self.ctx, ctypes.byref(n), None), "akz_descriptor_bytes")
        self.desc_bytes = n.value

    def close(self):
        if getattr(self, "ctx", None) and _lib is not None:
            _lib.akz_destroy(self.ctx)
        self.ctx = None

    def __del__(self):
        self.close()

    def input_buffer(self):
        """(device pointer of image 0, row pitch bytes, image stride bytes)."""
        p, pitch, zs = ctypes.c_void_p(), ctypes.c_size_t(), ctypes.c_size_t()
        _check(lib().akz_batch_input_buffer(self.ctx, ctypes.byref(p), ctypes.byref(pitch), ctypes.byref(zs)),
               "akz_batch_input_buffer")
        return p.value, pitch.value, zs.value

    def upload(self, imgs):
        """Copies host float32 images (n x H x W) into the context's input buffers."""
        ptr, pitch, zs = self.input_buffer()
        n = len(imgs)
        if n > self.batch:
            raise ValueError("batch capacity %d < %d images" % (self.batch, n))
        row = pitch // 4
        for z, img in enumerate(imgs):
            img = np.ascontiguousarray(img, np.float32)
            if img.shape != (self.H, self.W):
                raise ValueError("image shape does not match the options")
            padded = np.zeros((self.H, row), np.float32)
            padded[:, :self.W] = img
            _check(lib().akz_memcpy_h2d(ptr + z * zs, padded.ctypes.data, padded.nbytes), "akz_memcpy_h2d")
        return n

    def set_graph_mode(self, on):
        _check(lib().akz_set_graph_mode(self.ctx, int(bool(on))), "akz_set_graph_mode")

    def process_async(self, n, d_imgs=None, pitch_bytes=0, image_stride_bytes=0):
        """Runs images 0..n-1 (the context's input buffers unless d_imgs is given)."""
        if d_imgs is None:
            d_imgs, pitch_bytes, image_stride_bytes = self.input_buffer()
        _check(lib().akz_process_batch_device_async(self.ctx, d_imgs, pitch_bytes, image_stride_bytes, int(n)),
               "akz_process_batch_device_async")

    def counts(self, n):
        c = np.zeros(n, np.int32)
        st = np.zeros(n, np.int32)
        _check(lib().akz_read_batch_counts(self.ctx, int(n), c.ctypes.data, st.ctypes.data), "akz_read_batch_counts")
        return c, st

    def read_results(self, z, cap):
        kps = np.zeros(max(cap, 1), KP_DTYPE)
        desc = np.zeros((max(cap, 1), self.desc_bytes), np.uint8)
        n = ctypes.c_int()
        st = lib().akz_read_batch_results(self.ctx, int(z), kps.ctypes.data, desc.ctypes.data, cap, ctypes.byref(n))
        _check(st, "akz_read_batch_results", allow=(AKZ_ERR_OVERFLOW,))
        m = min(n.value, cap)
        return kps[:m].copy(), desc[:m].copy(), n.value

    def read_level(self, z, lvl, which):
        w, h = AKAZE.level_size(self, lvl)[:2]
        out = np.empty((h, w), np.float32)
        _check(lib().akz_read_batch_level(self.ctx, int(z), lvl, which, out.ctypes.data), "akz_read_batch_level")
        return out

    level_size = AKAZE.level_size

    def profile_fed(self, n):
        """(FED ms per call of n images, algorithmic bytes (12 B per pixel and
        step), launches, bytes the launches must move)."""
        d, pitch, zs = self.input_buffer()
        ms = ctypes.c_float(); by = ctypes.c_double(); k = ctypes.c_int(); mb = ctypes.c_double()
        _check(lib().akz_profile_fed_batch(self.ctx, d, pitch, zs, int(n), ctypes.byref(ms), ctypes.byref(by),
                                           ctypes.byref(k), ctypes.byref(mb)), "akz_profile_fed_batch")
        return ms.value, by.value, k.value, mb.value

    def wait(self):
        _check(lib().akz_wait(self.ctx), "akz_wait")

    def autotune(self, nz=None, reps=3):
        """Times the FED kernel variants per level size on this context's
        buffers (batches of nz images) and keeps the fastest."""
        _check(lib().akz_autotune(self.ctx, int(nz or self.batch), int(reps)), "akz_autotune")
        return self.fed_variants()

    def fed_variants(self):
        v = np.zeros(self.nlev, np.int32)
        _check(lib().akz_get_fed_variants(self.ctx, v.ctypes.data), "akz_get_fed_variants")
        return v

    def set_fed_variants(self, v):
        v = np.ascontiguousarray(v, np.int32)
        _check(lib().akz_set_fed_variants(self.ctx, v.ctypes.data), "akz_set_fed_variants")

    def set_keep_determinant(self, keep):
        """Write full determinant images in process_async runs (for read_level(.., 4)
        and extract); by default only the candidates' neighbourhoods are written."""
        _check(lib().akz_set_keep_determinant(self.ctx, int(bool(keep))), "akz_set_keep_determinant")

    def set_suppression_pool(self, entries):
        """Cap the suppression's LDS pool for long neighbour lists (0..4096);
        images whose lists do not fit take the global-memory path."""
        _check(lib().akz_set_suppression_pool(self.ctx, int(entries)), "akz_set_suppression_pool")

    def set_proposal(self, net, threshold=0.5):
        """Detect with the learned tile proposal `net` (None: full scan)."""
        _check(lib().akz_set_proposal(self.ctx, net.net if net is not None else None, float(threshold)),
               "akz_set_proposal")

    def proposal_positions(self):
        n = ctypes.c_int()
        _check(lib().akz_prop_positions(self.ctx, ctypes.byref(n)), "akz_prop_positions")
        return n.value

    def read_probs(self, z):
        out = np.zeros(self.proposal_positions(), np.float32)
        _check(lib().akz_prop_read_probs(self.ctx, int(z), out.ctypes.data), "akz_prop_read_probs")
        return out

    def results_device(self, z):
        """(device count pointer, keypoints pointer, 64-byte descriptor rows pointer) of image z."""
        c, k, d = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _check(lib().akz_batch_results_device(self.ctx, int(z), ctypes.byref(c), ctypes.byref(k), ctypes.byref(d)),
               "akz_batch_results_device")
        return c.value, k.value, d.value

    def extract(self, n, d_feat, d_labels=None):
        """Proposal features (and labels) of images 0..n-1 of the last run into device arrays."""
        _check(lib().akz_prop_extract(self.ctx, int(n), d_feat, d_labels), "akz_prop_extract")

    def synth(self, seed0, n):
        """Renders n device-generated synthetic images into the input buffers."""
        ptr, pitch, zs = self.input_buffer()
        _check(lib().akz_synth_device_batch(seed0, int(n), self.W, self.H, ptr, pitch, zs), "akz_synth_device_batch")

    def run(self, imgs):
        """Scale space + detection + description of host images; list of (kps, desc)."""
        n = self.upload(imgs)
        self.process_async(n)
        counts, _ = self.counts(n)
        return [self.read_results(z, int(counts[z]))[:2] for z in range(n)]


class Ragged:
    """Ragged batches: one batch context per size class, the classes on
    concurrent streams.  No counterpart in the reference."""

    def __init__(self, options, sizes, capacity):
        n = len(sizes)
        w = np.array([s[0] for s in sizes], np.int32)
        h = np.array([s[1] for s in sizes], np.int32)
        cap = np.array(capacity if np.iterable(capacity) else [capacity] * n, np.int32)
        p = ctypes.c_void_p()
        _check(lib().akz_ragged_create(ctypes.byref(options._c), n, w.ctypes.data, h.ctypes.data, cap.ctypes.data,
                                       ctypes.byref(p)), "akz_ragged_create")
        self.r = p.value
        self.sizes = [tuple(s) for s in sizes]
        self._bufs = []
        c, nb = ctypes.c_void_p(), ctypes.c_int()
        _check(lib().akz_ragged_context(self.r, 0, ctypes.byref(c)), "akz_ragged_context")
        _check(lib().akz_descriptor_bytes(c, ctypes.byref(nb), None), "akz_descriptor_bytes")
        self.desc_bytes = nb.value

    def close(self):
        if getattr(self, "r", None) and _lib is not None:
            _lib.akz_ragged_destroy(self.r)
        self.r = None

    def __del__(self):
        self.close()

    def set_graph_mode(self, on):
        _check(lib().akz_ragged_set_graph_mode(self.r, int(bool(on))), "akz_ragged_set_graph_mode")

    def set_work_lists(self, on):
        """Persistent detector / orientation / MLDB launches over the whole
        ragged call (default) or one batch call per size class."""
        _check(lib().akz_ragged_set_work_lists(self.r, int(bool(on))), "akz_ragged_set_work_lists")

    def autotune(self, reps=3):
        _check(lib().akz_ragged_autotune(self.r, int(reps)), "akz_ragged_autotune")

    def process_device(self, ptrs, pitches, sizes):
        """Device images (pointer, pitch bytes, (w, h)); waits; kept counts."""
        n = len(ptrs)
        P = (ctypes.c_void_p * n)(*ptrs)
        pit = np.array(pitches, np.uint64)
        w = np.array([s[0] for s in sizes], np.int32)
        h = np.array([s[1] for s in sizes], np.int32)
        counts = np.zeros(n, np.int32)
        _check(lib().akz_ragged_process(self.r, n, P, pit.ctypes.data, w.ctypes.data, h.ctypes.data,
                                        counts.ctypes.data), "akz_ragged_process")
        return counts

    def process(self, imgs):
        """Host float32 images of the class sizes; list of (kps, desc)."""
        self._bufs = []
        ptrs, pitches, sizes = [], [], []
        for img in imgs:
            img = np.ascontiguousarray(img, np.float32)
            b = DeviceBuffer(img.nbytes)
            b.upload(img)
            self._bufs.append(b)
            ptrs.append(b.ptr)
            pitches.append(img.strides[0])
            sizes.append((img.shape[1], img.shape[0]))
        counts = self.process_device(ptrs, pitches, sizes)
        return [self.read(i, int(c))[:2] for i, c in enumerate(counts)]

    def read(self, i, cap):
        kps = np.zeros(max(cap, 1), KP_DTYPE)
        desc = np.zeros((max(cap, 1), self.desc_bytes), np.uint8)
        n = ctypes.c_int()
        st = lib().akz_ragged_read(self.r, int(i), kps.ctypes.data, desc.ctypes.data, cap, ctypes.byref(n))
        _check(st, "akz_ragged_read", allow=(AKZ_ERR_OVERFLOW,))
        m = min(n.value, cap)
        return kps[:m].copy(), desc[:m].copy(), n.value

    def results_device(self, i):
        """(device count pointer, keypoints pointer, 64-byte descriptor rows pointer) of image i."""
        c, k, d = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _check(lib().akz_ragged_results_device(self.r, int(i), ctypes.byref(c), ctypes.byref(k), ctypes.byref(d)),
               "akz_ragged_results_device")
        return c.value, k.value, d.value


class Matcher:
    """libAKAZECU::Matcher (src/lib/AKAZE.h:29-55)."""

    def __init__(self):
        p = ctypes.c_void_p()
        _check(lib().akz_matcher_create(ctypes.byref(p)), "akz_matcher_create")
        self.m = p.value

    def __del__(self):
        if getattr(self, "m", None) and _lib is not None:
            _lib.akz_matcher_destroy(self.m)
        self.m = None

    def match2nn(self, q, t):
        """2-NN structured array (n_query x 2) of DMatch."""
        q = np.ascontiguousarray(q, np.uint8)
        t = np.ascontiguousarray(t, np.uint8)
        if q.ndim != 2 or t.ndim != 2 or q.shape[1] < 1 or q.shape[1] != t.shape[1]:
            raise ValueError("descriptors must be n x b uint8 with the same b (61 for 486-bit M-LDB)")
        if q.shape[1] > 61:
            t = np.ascontiguousarray(t[:, :61]); q = np.ascontiguousarray(q[:, :61])
        out = np.zeros(max(2 * len(q), 1), DM_DTYPE)
        _check(lib().akz_match2nn(self.m, q.ctypes.data, len(q), t.ctypes.data, len(t), q.shape[1],
                                  out.ctypes.data), "akz_match2nn")
        return out[:2 * len(q)].reshape(-1, 2)

    def bfmatch(self, q, t):
        """Matcher::bfmatch (AKAZE.cpp:36-79): n_query x 2 DMatch rows."""
        return self.match2nn(q, t)

    def bfmatch_(self, q, t):
        """n x 8 float32 [q, t1, 0, d1, q, t2, 0, d2] (AKAZE.cpp:82-105)."""
        m = self.match2nn(q, t)
        out = np.zeros((len(m), 8), np.float32)
        for k in range(2):
            out[:, 4 * k + 0] = m[:, k]["queryIdx"]
            out[:, 4 * k + 1] = m[:, k]["trainIdx"]
            out[:, 4 * k + 3] = m[:, k]["distance"]
        return out

    BFMatch = bfmatch_  # python_bindings.cpp:38

    def match2nn_device(self, d_q64, nq, d_t64, nt, d_out):
        _check(lib().akz_match2nn_device(self.m, d_q64, nq, d_t64, nt, d_out), "akz_match2nn_device")

    def wait(self):
        _check(lib().akz_matcher_wait(self.m), "akz_matcher_wait")

    def last_ms(self):
        ms = ctypes.c_float()
        _check(lib().akz_matcher_last_ms(self.m, ctypes.byref(ms)), "akz_matcher_last_ms")
        return ms.value
