"""Learned tile proposal: training and evaluation on the device.

The network (akaze_amd/csrc/akz_proposal.hip) predicts, per 16x16 tile of
every Hessian-response level, whether the tile holds a keypoint the plain
pipeline keeps.  A context with a network drops detector candidates outside
predicted tiles before the suppression.  Training data are device-rendered
synthetic images labelled by the plain pipeline itself (akz_prop_extract);
forward, backward and Adam run in the library's own kernels.
"""
import ctypes
import os
import time

import numpy as np

import akaze_amd as _pkg
from . import BatchAKAZE, AKAZEOptions, DeviceBuffer, lib, _check

NPARAM = 14513
WEIGHTS_DIR = os.path.dirname(os.path.abspath(__file__))


def num_params():
    n = ctypes.c_int()
    _check(lib().akz_prop_num_params(ctypes.byref(n)), "akz_prop_num_params")
    return n.value


class ProposalNet:
    def __init__(self, seed=1):
        p = ctypes.c_void_p()
        _check(lib().akz_prop_create(ctypes.c_uint64(seed), ctypes.byref(p)), "akz_prop_create")
        self.net = p.value

    def close(self):
        if getattr(self, "net", None):
            L = getattr(_pkg, "_lib", None)
            if L is not None:
                L.akz_prop_destroy(self.net)
        self.net = None

    def __del__(self):
        self.close()

    def get_params(self):
        out = np.zeros(NPARAM, np.float32)
        _check(lib().akz_prop_get_params(self.net, out.ctypes.data), "akz_prop_get_params")
        return out

    def set_params(self, params):
        params = np.ascontiguousarray(params, np.float32)
        assert params.shape == (NPARAM,)
        _check(lib().akz_prop_set_params(self.net, params.ctypes.data), "akz_prop_set_params")

    def save(self, path):
        self.get_params().astype("<f4").tofile(path)

    def load(self, path):
        self.set_params(np.fromfile(path, "<f4"))

    def forward(self, ctx, d_feat, nimg, idx=None):
        P = ctx.proposal_positions()
        out = np.zeros(nimg * P, np.float32)
        ip = None
        if idx is not None:
            idx = np.ascontiguousarray(idx, np.int32)
            ip = idx.ctypes.data
        _check(lib().akz_prop_forward(self.net, ctx.ctx, d_feat, ip, int(nimg), out.ctypes.data), "akz_prop_forward")
        return out

    def train_step(self, ctx, d_feat, d_labels, nimg, idx=None, pos_weight=1.0, lr=1e-3, want_grad=False):
        g = np.zeros(NPARAM, np.float32) if want_grad else None
        loss = ctypes.c_float()
        ip = None
        if idx is not None:
            idx = np.ascontiguousarray(idx, np.int32)
            ip = idx.ctypes.data
        _check(lib().akz_prop_train_step(self.net, ctx.ctx, d_feat, d_labels, ip, int(nimg), float(pos_weight),
                                         float(lr), g.ctypes.data if g is not None else None, ctypes.byref(loss)),
               "akz_prop_train_step")
        return loss.value, g


class Dataset:
    """Features and labels of n device-rendered images, resident in HBM."""

    def __init__(self, W, H, n, seed0, batch=16, log=None):
        self.W, self.H, self.n, self.seed0 = W, H, n, seed0
        self.ctx = BatchAKAZE(AKAZEOptions(img_width=W, img_height=H), batch)
        self.ctx.set_keep_determinant(True)  # the features read the determinant images
        self.P = self.ctx.proposal_positions()
        self.feat = DeviceBuffer(n * self.P * 8)
        self.lab = DeviceBuffer(n * self.P)
        t0 = time.perf_counter()
        for s in range(0, n, batch):
            k = min(batch, n - s)
            self.ctx.synth(seed0 + s, k)
            self.ctx.process_async(k)
            self.ctx.extract(k, self.feat.ptr + s * self.P * 8, self.lab.ptr + s * self.P)
            if log and (s // batch) % 50 == 0:
                log("dataset %dx%d: %d/%d images (%.1f s)" % (W, H, s + k, n, time.perf_counter() - t0))
        self.labels = np.zeros(n * self.P, np.uint8)
        self.lab.download(self.labels)
        self.pos_frac = float(self.labels.mean())


def train(net, data, steps, batch=32, lr=1e-3, pos_weight=None, seed=0, log=None, log_every=100):
    """Adam on random image batches of `data`; returns the loss history."""
    rng = np.random.default_rng(seed)
    if pos_weight is None:
        pos_weight = 1.0
    hist = []
    order = rng.permutation(data.n)
    pos = 0
    t0 = time.perf_counter()
    for it in range(steps):
        if pos + batch > data.n:
            order = rng.permutation(data.n)
            pos = 0
        idx = order[pos:pos + batch]
        pos += batch
        loss, _ = net.train_step(data.ctx, data.feat.ptr, data.lab.ptr, len(idx), idx=idx, pos_weight=pos_weight,
                                 lr=lr)
        hist.append(loss)
        if log and (it % log_every == 0 or it == steps - 1):
            log("step %d/%d loss %.5f (%.1f s)" % (it + 1, steps, float(np.mean(hist[-log_every:])),
                                                     time.perf_counter() - t0))
    return hist


def tile_recall(net, data, threshold, nimg=None):
    """Fraction of positive tiles with probability >= threshold, and the
    fraction of all tiles kept, over the first nimg images of `data`."""
    nimg = nimg or data.n
    probs = []
    for s in range(0, nimg, 64):
        k = min(64, nimg - s)
        probs.append(net.forward(data.ctx, data.feat.ptr, k, idx=np.arange(s, s + k)))
    prob = np.concatenate(probs)
    lab = data.labels[:nimg * data.P].astype(bool)
    keep = prob >= threshold
    return float((keep & lab).sum() / max(lab.sum(), 1)), float(keep.mean()), prob


def threshold_for_recall(prob, labels, recall):
    """Largest threshold whose tile recall on (prob, labels) is >= recall."""
    pos = np.sort(prob[labels.astype(bool)])
    if len(pos) == 0:
        return 0.5
    k = int(np.floor((1.0 - recall) * len(pos)))
    return float(np.nextafter(pos[k], -np.inf)) if k < len(pos) else 0.0


def keypoint_recall(full, prop):
    """Fraction of the full scan's keypoints (x, y, level, response) that the
    proposal run reproduces exactly."""
    if len(full) == 0:
        return 1.0
    key = lambda k: set(zip(k["x"].tolist(), k["y"].tolist(), k["class_id"].tolist(), k["response"].tolist()))
    return len(key(full) & key(prop)) / len(full)
