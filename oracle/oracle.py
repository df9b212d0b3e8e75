"""ctypes front end of the CPU oracle (oracle/sso_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the library.  ``Oracle('f32')`` is the
fp32 restatement of the reference's per-frame path in its operation order;
``Oracle('f64')`` evaluates the same method in fp64 and is the yardstick the
fp32 (and later fp16/bf16 volume) results are measured against.

Every method takes and returns numpy arrays in the reference's layouts
(HWC images, planar [N][H][W] volumes, uint8 [H][W][3] guides).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_f = C.c_int
_d = C.c_double
_p = C.c_void_p


def build():
    """Compiles both oracle libraries with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _lib_path(precision):
    return os.path.join(HERE, "lib", "liboracle_%s.so" % precision)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    def __init__(self, precision="f32"):
        assert precision in ("f32", "f64")
        path = _lib_path(precision)
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        self.sfx = "_" + precision
        self.real = np.float32 if precision == "f32" else np.float64
        self.precision = precision
        fn = self._fn("ora_expf")
        fn.restype = C.c_float if precision == "f32" else C.c_double
        fn.argtypes = [C.c_float if precision == "f32" else C.c_double]
        self._levels = self._fn("ora_isb_levels")
        self._levels.restype = C.c_int
        for name, args in _SIGNATURES.items():
            f = self._fn(name)
            f.argtypes = args
            f.restype = C.c_int if name == "ora_isb_filter" else None

    def _fn(self, name):
        return getattr(self.lib, name + self.sfx)

    @staticmethod
    def _d64(a):
        return np.ascontiguousarray(a, dtype=np.float64)

    # ---- math -------------------------------------------------------------
    def expf(self, x):
        fn = self._fn("ora_expf")
        return np.array([fn(float(v)) for v in np.asarray(x, dtype=np.float64).ravel()], dtype=self.real)

    def levels(self, W, H):
        return self._levels(_f(W), _f(H))

    # ---- double-sphere model -----------------------------------------------
    def unproject_torch(self, cam6, uv):
        uv = self._d64(uv).reshape(-1, 2)
        out = np.zeros((uv.shape[0], 3), self.real)
        valid = np.zeros(uv.shape[0], np.int32)
        self._fn("ora_unproject_torch")(_ptr(self._d64(cam6)), _ptr(uv), _f(uv.shape[0]), _ptr(out), _ptr(valid))
        return out, valid.astype(bool)

    def project_torch(self, cam6, pts):
        pts = self._d64(pts).reshape(-1, 3)
        uv = np.zeros((pts.shape[0], 2), self.real)
        valid = np.zeros(pts.shape[0], np.int32)
        self._fn("ora_project_torch")(_ptr(self._d64(cam6)), _ptr(pts), _f(pts.shape[0]), _ptr(uv), _ptr(valid))
        return uv, valid.astype(bool)

    def unproject_cuda(self, intr6, uv):
        uv = self._d64(uv).reshape(-1, 2)
        out = np.zeros((uv.shape[0], 3), self.real)
        self._fn("ora_unproject_cuda")(_ptr(self._d64(intr6)), _ptr(uv), _f(uv.shape[0]), _ptr(out))
        return out

    def project_cuda(self, intr6, pts, with_valid=False):
        pts = self._d64(pts).reshape(-1, 3)
        uv = np.zeros((pts.shape[0], 2), self.real)
        valid = np.zeros(pts.shape[0], np.int32)
        self._fn("ora_project_cuda")(_ptr(self._d64(intr6)), _ptr(pts), _f(pts.shape[0]), _ptr(uv),
                                     _ptr(valid) if with_valid else None)
        return (uv, valid.astype(bool)) if with_valid else uv

    # ---- estimator stages ---------------------------------------------------
    def select_camera(self, cams6, rts, masks, ref, min_dist, max_dist):
        masks = np.ascontiguousarray(masks, np.float32)
        Cn, H, W = masks.shape
        sel = np.zeros((H, W), np.int32)
        self._fn("ora_select_camera")(_f(Cn), _f(H), _f(W), _f(ref), _ptr(self._d64(cams6)), _ptr(self._d64(rts)),
                                      _ptr(masks), _d(min_dist), _d(max_dist), _ptr(sel))
        return sel

    def sweep(self, imgs, sel, cands, cams6, rts, ref, vtype=0):
        """Cost volume [N][H][W]; vtype 1 / 2 rounds every stored value to
        fp16 / bf16 (the precision form; ignored by the fp64 oracle)."""
        imgs = np.ascontiguousarray(imgs, np.float32)
        Cn, H, W, _ = imgs.shape
        sel = np.ascontiguousarray(sel, np.int32)
        cands = self._d64(cands)
        N = cands.shape[0]
        cost = np.zeros((N, H, W), self.real)
        self._fn("ora_sweep")(_f(Cn), _f(H), _f(W), _f(ref), _ptr(imgs), _ptr(sel), _ptr(cands), _f(N),
                              _ptr(self._d64(cams6)), _ptr(self._d64(rts)), _ptr(cost), _f(vtype))
        return cost

    def guide(self, img):
        img = np.ascontiguousarray(img, np.float32)
        H, W, _ = img.shape
        g = np.zeros((H, W, 3), np.uint8)
        self._fn("ora_guide")(_ptr(img), _f(H * W), _ptr(g))
        return g

    def isb_down(self, gin, cin, hout, wout, var_inv_i, vtype=0):
        gin = np.ascontiguousarray(gin, np.uint8)
        cin = np.ascontiguousarray(cin, self.real)
        N, hin, win = cin.shape
        gout = np.zeros((hout, wout, 3), np.uint8)
        cout = np.zeros((N, hout, wout), self.real)
        self._fn("ora_isb_down")(_ptr(gin), _ptr(cin), _f(hin), _f(win), _ptr(gout), _ptr(cout), _f(hout),
                                 _f(wout), _f(N), _d(var_inv_i), _f(vtype))
        return gout, cout

    def isb_up(self, gin, cin, gio, cio, weight_up, weight_down, var_inv_i, vtype=0):
        """Returns updated copies of (gio, cio)."""
        gin = np.ascontiguousarray(gin, np.uint8)
        cin = np.ascontiguousarray(cin, self.real)
        gio = np.array(gio, np.uint8, copy=True, order="C")
        cio = np.array(cio, self.real, copy=True, order="C")
        N, hin, win = cin.shape
        _, hio, wio = cio.shape
        self._fn("ora_isb_up")(_ptr(gin), _ptr(cin), _f(hin), _f(win), _ptr(gio), _ptr(cio), _f(hio), _f(wio),
                               _f(N), _d(weight_up), _d(weight_down), _d(var_inv_i), _f(vtype))
        return gio, cio

    def isb_filter(self, guide, cost, sigma_i, sigma_s, vtype=0, round_level0=True):
        """ISB_Filter.apply on copies: returns (cost, guide) filtered.  vtype:
        storage rounding of every level written; round_level0=False leaves
        the final level-0 values unrounded (the fused argmin reads them)."""
        guide = np.array(guide, np.uint8, copy=True, order="C")
        cost = np.array(cost, self.real, copy=True, order="C")
        if cost.ndim == 2:
            cost = cost[None]
        N, H, W = cost.shape
        rc = self._fn("ora_isb_filter")(_ptr(guide), _ptr(cost), _f(H), _f(W), _f(N), _d(sigma_i), _d(sigma_s),
                                        _f(vtype), _f(int(bool(round_level0))))
        if rc < 0:
            raise ValueError("resolution %dx%d gives no pyramid level" % (W, H))
        return cost, guide

    def argmin_refine(self, cost, d0, dN1, A):
        cost = np.ascontiguousarray(cost, self.real)
        N, H, W = cost.shape
        dist = np.zeros((H, W), self.real)
        idx = np.zeros((H, W), np.int32)
        self._fn("ora_argmin_refine")(_ptr(cost), _f(N), _f(H * W), _d(d0), _d(dN1), _d(A), _ptr(dist), _ptr(idx))
        return dist, idx

    def soft_argmin(self, cost, beta, inv0, inv1):
        """The differentiation form's readout: inverse distance per pixel of a
        planar [N][H][W] filtered volume."""
        cost = np.ascontiguousarray(cost, self.real)
        N, H, W = cost.shape
        invd = np.zeros((H, W), self.real)
        self._fn("ora_soft_argmin")(_ptr(cost), _f(N), _f(H * W), _d(beta), _d(inv0), _d(inv1), _ptr(invd))
        return invd

    def round_volume(self, x, vtype):
        """Values rounded as the precision form stores them (1 fp16, 2 bf16)."""
        x = np.array(x, self.real, copy=True, order="C")
        self._fn("ora_round_volume")(_ptr(x), C.c_long(x.size), _f(vtype))
        return x

    # ---- stitcher ------------------------------------------------------------
    def reproject_zmin(self, din, intr6, t3):
        din = np.ascontiguousarray(din, self.real)
        H, W = din.shape
        dout = np.full((H, W), 1e8, self.real)
        self._fn("ora_reproject_zmin")(_ptr(din), _ptr(dout), _f(H), _f(W), _ptr(self._d64(intr6)),
                                       _ptr(self._d64(t3)))
        return dout

    def inpaint_lut(self, intr6, t3, H, W, min_dist, max_dist):
        lut = np.zeros((H, W, 2), np.uint8)
        self._fn("ora_inpaint_lut")(_ptr(lut), _f(H), _f(W), _ptr(self._d64(intr6)), _ptr(self._d64(t3)),
                                    _d(min_dist), _d(max_dist))
        return lut

    def inpaint_step(self, din, lut, thr_empty, thr_src):
        din = np.ascontiguousarray(din, self.real)
        H, W = din.shape
        dout = np.zeros_like(din)
        self._fn("ora_inpaint_step")(_ptr(din), _ptr(dout), _ptr(np.ascontiguousarray(lut, np.uint8)), _f(H), _f(W),
                                     _d(thr_empty), _d(thr_src))
        return dout

    def resize_linear(self, img, cols, rows):
        """The input stage's resize (utils.py:198-201) of an HWC or HW image
        already converted to float32 in [0, 255]."""
        img = np.ascontiguousarray(img, np.float32)
        hw = img.ndim == 2
        x = img[..., None] if hw else img
        Hs, Ws, Cn = x.shape
        out = np.zeros((rows, cols, Cn), self.real)
        self._fn("ora_resize_linear")(_ptr(x), _f(Hs), _f(Ws), _f(Cn), _ptr(out), _f(rows), _f(cols))
        return out[..., 0] if hw else out

    def inpaint(self, d, lut, iterations, thr_empty, thr_src):
        for _ in range(iterations):
            d = self.inpaint_step(d, lut, thr_empty, thr_src)
        return d

    def box_filter(self, x, radius, pad_value):
        x = np.ascontiguousarray(x, self.real)
        squeeze = x.ndim == 2
        if squeeze:
            x = x[None]
        B, H, W = x.shape
        out = np.zeros_like(x)
        self._fn("ora_box_filter")(_ptr(x), _ptr(out), _f(B), _f(H), _f(W), _f(radius), _d(pad_value))
        return out[0] if squeeze else out

    def blend_lut(self, intr, rot, trans, masks, PH, PW, min_dist, max_dist):
        masks = np.ascontiguousarray(masks, self.real)
        R, H, W = masks.shape
        sampling = np.zeros((R, PH, PW, 2), self.real)
        weights = np.zeros((R, PH, PW), self.real)
        self._fn("ora_blend_lut")(_ptr(sampling), _ptr(weights), _ptr(masks), _f(R), _f(H), _f(W), _f(PH), _f(PW),
                                  _ptr(self._d64(intr)), _ptr(self._d64(rot)), _ptr(self._d64(trans)), _d(min_dist),
                                  _d(max_dist))
        return sampling, weights

    def normalize_weights(self, w):
        w = np.array(w, self.real, copy=True, order="C")
        R = w.shape[0]
        self._fn("ora_normalize_weights")(_ptr(w), _f(R), _f(w[0].size))
        return w

    def merge(self, sampling, weights, reproj, dists, imgs, trans, intr):
        sampling = np.ascontiguousarray(sampling, self.real)
        weights = np.ascontiguousarray(weights, self.real)
        reproj = np.ascontiguousarray(reproj, self.real)
        dists = np.ascontiguousarray(dists, self.real)
        imgs = np.ascontiguousarray(imgs, np.uint8)
        R, PH, PW = weights.shape
        _, H, W = reproj.shape
        _, Hs, Ws, _ = imgs.shape
        dist = np.zeros((PH, PW), self.real)
        rgb = np.zeros((PH, PW, 3), np.uint8)
        self._fn("ora_merge")(_ptr(sampling), _ptr(weights), _ptr(reproj), _ptr(dists), _ptr(imgs), _f(Hs), _f(Ws),
                              _ptr(self._d64(trans)), _ptr(self._d64(intr)), _f(R), _f(H), _f(W), _f(PH), _f(PW),
                              _ptr(dist), _ptr(rgb))
        return dist, rgb


_SIGNATURES = {
    "ora_unproject_torch": [_p, _p, _f, _p, _p],
    "ora_project_torch": [_p, _p, _f, _p, _p],
    "ora_unproject_cuda": [_p, _p, _f, _p],
    "ora_project_cuda": [_p, _p, _f, _p, _p],
    "ora_select_camera": [_f, _f, _f, _f, _p, _p, _p, _d, _d, _p],
    "ora_sweep": [_f, _f, _f, _f, _p, _p, _p, _f, _p, _p, _p, _f],
    "ora_guide": [_p, _f, _p],
    "ora_isb_down": [_p, _p, _f, _f, _p, _p, _f, _f, _f, _d, _f],
    "ora_isb_up": [_p, _p, _f, _f, _p, _p, _f, _f, _f, _d, _d, _d, _f],
    "ora_isb_filter": [_p, _p, _f, _f, _f, _d, _d, _f, _f],
    "ora_round_volume": [_p, C.c_long, _f],
    "ora_argmin_refine": [_p, _f, _f, _d, _d, _d, _p, _p],
    "ora_soft_argmin": [_p, _f, _f, _d, _d, _d, _p],
    "ora_reproject_zmin": [_p, _p, _f, _f, _p, _p],
    "ora_inpaint_lut": [_p, _f, _f, _p, _p, _d, _d],
    "ora_inpaint_step": [_p, _p, _p, _f, _f, _d, _d],
    "ora_box_filter": [_p, _p, _f, _f, _f, _f, _d],
    "ora_blend_lut": [_p, _p, _p, _f, _f, _f, _f, _f, _p, _p, _p, _d, _d],
    "ora_normalize_weights": [_p, _f, _f],
    "ora_merge": [_p, _p, _p, _p, _p, _f, _f, _p, _p, _f, _f, _f, _f, _f, _p, _p],
    "ora_resize_linear": [_p, _f, _f, _f, _p, _f, _f],
}
