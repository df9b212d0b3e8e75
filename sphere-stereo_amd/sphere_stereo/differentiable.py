"""Differentiation form: a soft-argmin depth readout with hand-written backward
kernels, and a photometric Gauss-Newton refinement of a camera's calibration.

The forward of one reference camera is the plain path's own stages on an fp32
volume (unit rays, sweep + clamped cost, ISB cost filter; depth_estimation.py:
141-177), read out by a soft-argmin over the filtered volume instead of the
hard argmin + parabola of depth_estimation.py:180-198:

    iota = sum_n n softmax_n(-F_n / (cost_scale * temperature)),
    inverse distance = 1/min_dist + (1/max_dist - 1/min_dist) * iota / (N - 1).

``DifferentiableDepth`` exposes it as a ``torch.autograd.Function`` whose
backward runs the adjoint kernels (soft-argmin, ISB filter, sweep) and returns
gradients for the matching images, the double-sphere parameters of every
camera (fs0, fs1, cs0, cs1, xi, alpha in matching-resolution units) and the
camera-to-rig poses (4 x 4, through the relative transforms, whose small
4 x 4 algebra stays in torch float64).  No pyramid level of the forward
volume is stored: the adjoint recomputes the guide pyramid.  The camera
selection (depth_estimation.py:83-135), the uint8 guide and the bilinear cell
are piecewise constant and contribute no derivative.

``refine_calibration`` uses the forward to estimate the references' depth and
the Gauss-Newton kernels to fit one camera's pose and focal lengths to the
photometric reprojection of the references' pixels, the small normal
equations being solved in fp64 on the GPU.
"""
import ctypes as C
import math

import torch

from . import _lib
from .utils import distance_candidates


def _cam_struct(row):
    """ss_camera from (fs0, fs1, cs0, cs1, xi, alpha)."""
    r = [float(v) for v in row]
    c = _lib.ss_camera()
    c.fs[0], c.fs[1], c.cs[0], c.cs[1] = r[0], r[1], r[2], r[3]
    c.xi, c.alpha = r[4], r[5]
    return c


def _cam_array(cams6):
    arr = (_lib.ss_camera * cams6.shape[0])()
    for i, row in enumerate(cams6.tolist()):
        arr[i] = _cam_struct(row)
    return arr


def camera_parameters(calibrations, pixel_centres="reference"):
    """[C, 6] float64: focal lengths, principal point, xi, alpha at the
    matching resolution.  Focal lengths are fl * matching_scale.  The
    principal point with pixel_centres = "reference" is principal *
    matching_scale (the fp32 products the reference forms, utils.py:62, 94);
    with "resized" it is (principal + 0.5) * matching_scale - 0.5, where the
    original image's pixel centres land in an image resized with cv2's
    INTER_LINEAR mapping (src = (dst + 0.5) / scale - 0.5, utils.py:198-201).
    The two differ by 0.5 * (1 - scale) px (0.18 px from 1216^2 to 768^2)."""
    if pixel_centres not in ("reference", "resized"):
        raise ValueError("pixel_centres must be 'reference' or 'resized'")
    rows = []
    for c in calibrations:
        fl = (c.fl.float().cpu() * c.matching_scale.float().cpu()).tolist()
        if pixel_centres == "reference":
            pp = (c.principal.float().cpu() * c.matching_scale.float().cpu()).tolist()
        else:
            pp = ((c.principal.double().cpu() + 0.5) * c.matching_scale.double().cpu() - 0.5).tolist()
        rows.append(fl + pp + [float(c.xi), float(c.alpha)])
    return torch.tensor(rows, dtype=torch.float64)


def camera_poses(calibrations):
    """[C, 4, 4] float64 camera-to-rig poses."""
    return torch.stack([c.rt.detach().double().cpu() for c in calibrations])


def relative_poses(rts, ref):
    """[C, 3, 4]: rows 0-2 of inv(rt_k) @ rt_ref (camera k <- reference), the
    transform of depth_estimation.py:156, differentiable in float64."""
    return (torch.linalg.inv(rts) @ rts[ref])[:, :3, :]


def _skew(w):
    z = torch.zeros_like(w[0])
    return torch.stack([torch.stack([z, -w[2], w[1]]), torch.stack([w[2], z, -w[0]]),
                        torch.stack([-w[1], w[0], z])])


def pose_exp(delta):
    """4 x 4 float64 transform [[exp(omega^), v], [0, 1]] of a 6-vector (omega,
    v), differentiable everywhere (series near omega = 0)."""
    delta = delta.to(torch.float64)
    w, v = delta[:3], delta[3:6]
    t2 = (w * w).sum()
    small = t2 < 1e-8
    t2s = torch.where(small, torch.ones_like(t2), t2)
    th = torch.sqrt(t2s)
    a = torch.where(small, 1 - t2 / 6, torch.sin(th) / th)
    b = torch.where(small, 0.5 - t2 / 24, (1 - torch.cos(th)) / t2s)
    K = _skew(w)
    R = torch.eye(3, dtype=torch.float64, device=delta.device) + a * K + b * (K @ K)
    top = torch.cat([R, v[:, None]], dim=1)
    bottom = torch.tensor([[0.0, 0.0, 0.0, 1.0]], dtype=torch.float64, device=delta.device)
    return torch.cat([top, bottom], dim=0)


class _SoftDepth(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rig, slot, cams6, rel, *images):
        invd, state = rig._forward_slot(slot, cams6.detach(), rel.detach(), images)
        ctx.rig, ctx.slot, ctx.state = rig, slot, state
        ctx.img_grad = [im.requires_grad for im in images]
        ctx.save_for_backward(cams6, rel, *images)
        return invd

    @staticmethod
    def backward(ctx, g):
        cams6, rel, *images = ctx.saved_tensors
        g_imgs, g_cams, g_rel = ctx.rig._backward_slot(ctx.slot, cams6.detach(), rel.detach(), images,
                                                        g.contiguous(), ctx.state, any(ctx.img_grad))
        ctx.state = None
        g_imgs = [gi if need else None for gi, need in zip(g_imgs, ctx.img_grad)]
        return (None, None, g_cams.to(cams6.dtype), g_rel.to(rel.dtype), *g_imgs)


class DifferentiableDepth:
    """Soft-argmin inverse-distance maps of a rig's reference cameras,
    differentiable with respect to the images, intrinsics and extrinsics.

    Arguments as RGBD_Estimator's (depth_estimation.py:39-40) for the matching
    part, plus ``temperature`` of the soft-argmin and ``cost_scale``, the
    cost the temperature is relative to (the clamp value 500 of
    depth_estimation.py:175).  The cost volume is fp32.

    ``pixel_centres`` (camera_parameters): "resized" (default) places the
    principal point where the calibration's pixel centres land in images
    resized to the matching resolution with INTER_LINEAR, as read_input_images
    resizes them; "reference" keeps the reference's principal * scale, which
    is 0.5 * (1 - scale) px off for resized images.  A calibration refined
    against the images converges to the former: with the latter as the
    starting truth, the refinement of a 768^2 rig settled ~3 mm and 0.02
    degrees away from it, and 0.3 mm and 0.002 degrees with the former.
    """

    def __init__(self, calibrations, min_dist, max_dist, candidate_count, references_indices, masks,
                 matching_resolution, sigma_i, sigma_s, device, temperature=0.05, cost_scale=500.0,
                 pixel_centres="resized"):
        self.device = torch.device(device)
        self.C = len(calibrations)
        self.refs = list(references_indices)
        self.N = int(candidate_count)
        if self.N < 2 or self.N > _lib.SS_MAX_CANDIDATES:
            raise ValueError("candidate_count must be in [2, %d]" % _lib.SS_MAX_CANDIDATES)
        if not 1 <= self.C <= _lib.SS_MAX_CAMERAS:
            raise ValueError("between 1 and %d cameras supported" % _lib.SS_MAX_CAMERAS)
        self.cols, self.rows = int(matching_resolution[0]), int(matching_resolution[1])
        self.min_dist, self.max_dist = float(min_dist), float(max_dist)
        self.sigma_i, self.sigma_s = float(sigma_i), float(sigma_s)
        self.beta = 1.0 / (float(cost_scale) * float(temperature))
        self.inv0, self.inv1 = 1.0 / self.min_dist, 1.0 / self.max_dist
        self.NP = _lib.lib.ss_padded_candidates(self.N)
        self._cands = _lib.float_array(distance_candidates(min_dist, max_dist, self.N))
        self.cams6 = camera_parameters(calibrations, pixel_centres)
        self.rts = camera_poses(calibrations)
        self.masks = torch.stack([m.reshape(self.rows, self.cols).float() for m in masks]).to(self.device).contiguous()
        H, W, R = self.rows, self.cols, len(self.refs)
        self._sel = [torch.empty(H * W, dtype=torch.int8, device=self.device) for _ in range(R)]
        nb = _lib.lib.ss_isb_workspace_bytes(H, W, self.N, _lib.VOLUME_TYPES["float32"], 0)
        nbb = _lib.lib.ss_isb_backward_workspace_bytes(H, W, self.N)
        nbs = _lib.lib.ss_sweep_backward_workspace_bytes(H, W, self.C)
        if nb == 0 or nbb == 0 or nbs == 0:
            raise ValueError("unsupported matching resolution / candidate count")
        self._ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
        self._ws_b = torch.empty(nbb, dtype=torch.uint8, device=self.device)
        self._ws_s = torch.empty(nbs, dtype=torch.uint8, device=self.device)
        self.set_calibration(self.cams6, self.rts)

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def set_calibration(self, cams6, rts):
        """Recomputes the camera selection of every reference
        (depth_estimation.py:83-135) for this calibration; the selection is
        then held fixed by the forward and its gradient."""
        cams6 = cams6.detach().double().cpu()
        rts = rts.detach().double().cpu()
        arr = _cam_array(cams6)
        for slot, ref in enumerate(self.refs):
            rel = relative_poses(rts, ref).reshape(-1).tolist()
            _lib.check(_lib.lib.ss_select_camera(arr, _lib.float_array(rel), self.C, ref, self.masks.data_ptr(),
                                                 self.rows, self.cols, self.min_dist, self.max_dist,
                                                 self._sel[slot].data_ptr(), self._stream()))

    def _forward_slot(self, slot, cams6, rel, images):
        H, W, N = self.rows, self.cols, self.N
        ref = self.refs[slot]
        imgs = [im.contiguous() for im in images]
        for im in imgs:
            if im.dtype != torch.float32 or tuple(im.shape) != (H, W, 3) or im.device != self.device:
                raise ValueError("images must be float32 [%d, %d, 3] on %s" % (H, W, self.device))
        arr = _cam_array(cams6.cpu())
        rts = _lib.float_array(rel.cpu().reshape(-1).tolist())
        s = self._stream()
        raylut = torch.empty((H, W, 4), dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib.ss_ray_lut(C.byref(arr[ref]), self._sel[slot].data_ptr(), H, W, raylut.data_ptr(), s))
        vol = torch.empty(H * W * self.NP, dtype=torch.float32, device=self.device)
        guide = torch.empty(H * W, dtype=torch.int32, device=self.device)
        _lib.check(_lib.lib.ss_sweep(raylut.data_ptr(), _lib.ptr_array(imgs), self.C, arr, rts, ref, self._cands, N,
                                     H, W, _lib.VOLUME_TYPES["float32"], vol.data_ptr(), guide.data_ptr(), s))
        guide0 = guide.clone()
        _lib.check(_lib.lib.ss_isb_filter(guide.data_ptr(), vol.data_ptr(), H, W, N, _lib.VOLUME_TYPES["float32"],
                                          self.sigma_i, self.sigma_s, self._ws.data_ptr(), self._ws.numel(), s))
        invd = torch.empty((H, W), dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib.ss_soft_argmin(vol.data_ptr(), H, W, N, self.beta, self.inv0, self.inv1, invd.data_ptr(),
                                           s))
        return invd, (raylut, vol, guide0, imgs)

    def _backward_slot(self, slot, cams6, rel, images, g, state, want_images):
        H, W, N, Cn = self.rows, self.cols, self.N, self.C
        ref = self.refs[slot]
        raylut, vol, guide0, imgs = state
        s = self._stream()
        gF = torch.empty_like(vol)
        _lib.check(_lib.lib.ss_soft_argmin_backward(vol.data_ptr(), g.data_ptr(), H, W, N, self.beta, self.inv0,
                                                    self.inv1, gF.data_ptr(), s))
        gC = torch.empty_like(vol)
        _lib.check(_lib.lib.ss_isb_filter_backward(guide0.data_ptr(), gF.data_ptr(), gC.data_ptr(), H, W, N,
                                                   self.sigma_i, self.sigma_s, self._ws_b.data_ptr(),
                                                   self._ws_b.numel(), s))
        del gF
        g_imgs = [torch.zeros_like(im) for im in imgs] if want_images else None
        gp = torch.empty((Cn + 1) * 18, dtype=torch.float32, device=self.device)
        arr = _cam_array(cams6.cpu())
        rts = _lib.float_array(rel.cpu().reshape(-1).tolist())
        _lib.check(_lib.lib.ss_sweep_backward(raylut.data_ptr(), _lib.ptr_array(imgs), Cn, arr, rts, ref, self._cands,
                                              N, H, W, gC.data_ptr(),
                                              _lib.ptr_array(g_imgs) if g_imgs is not None else None, gp.data_ptr(),
                                              self._ws_s.data_ptr(), self._ws_s.numel(), s))
        gp = gp.view(Cn + 1, 18).double().cpu()
        g_cams = gp[:Cn, :6].clone()
        g_cams[ref] += gp[Cn, :6]
        g_rel = gp[:Cn, 6:].reshape(Cn, 3, 4).clone()
        if g_imgs is None:
            g_imgs = [None] * Cn
        return g_imgs, g_cams.to(cams6.device), g_rel.to(rel.device)

    def _hard_distance(self, vol):
        """Argmin + parabola refinement + index to distance
        (depth_estimation.py:180-198) of a filtered fp32 volume."""
        d0, dN1 = float(self._cands[0]), float(self._cands[self.N - 1])
        A = float(torch.tensor(d0, dtype=torch.float32) / torch.tensor(dN1, dtype=torch.float32) - 1)
        dist = torch.empty((self.rows, self.cols), dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib.ss_argmin_refine(vol.data_ptr(), self.rows, self.cols, self.N,
                                             _lib.VOLUME_TYPES["float32"], d0, dN1, A, dist.data_ptr(),
                                             self._stream()))
        return dist

    def __call__(self, images, cams6=None, rts=None):
        """Soft-argmin inverse-distance maps [rows, cols] of the references for
        the C matching images (float32 [rows, cols, 3] in [0, 255]).  cams6
        ([C, 6] float64) and rts ([C, 4, 4] float64, camera to rig) default to
        the constructor's calibration; gradients flow to whichever of images,
        cams6 and rts require them."""
        cams6 = self.cams6 if cams6 is None else cams6
        rts = self.rts if rts is None else rts
        return [_SoftDepth.apply(self, slot, cams6, relative_poses(rts, ref), *images)
                for slot, ref in enumerate(self.refs)]

    # ---- photometric calibration refinement ---------------------------------------
    def refine_calibration(self, frames, target, steps=50, huber=10.0, damping=1e-3, cams6=None, rts=None,
                           depth="hard", skip_selected=True):
        """Gauss-Newton fit of camera ``target``'s pose (a right perturbation of
        its camera-to-rig transform) and focal lengths (fs0, fs1 scales) to
        the photometric reprojection of the references' pixels.  Each step
        re-estimates the references' depth with the current calibration
        (depth = "hard": the argmin with parabola refinement of
        depth_estimation.py:180-198 over the filtered volume; "soft": the
        soft-argmin readout), skips the pixels whose depth was matched
        against the target itself, accumulates the normal equations of all
        frames (lists of C matching images of one rig) and solves them in
        fp64 on the GPU.  Returns (cams6, rts, history): history holds, per
        step, the mean Huber loss and residual count before the step and the
        target's parameters after it."""
        if depth not in ("hard", "soft"):
            raise ValueError("depth must be 'hard' or 'soft'")
        if target in self.refs:
            raise ValueError("the refined camera must not be a reference")
        cams6 = (self.cams6 if cams6 is None else cams6).detach().double().cpu().clone()
        rts = (self.rts if rts is None else rts).detach().double().cpu().clone()
        H, W = self.rows, self.cols
        nb = _lib.lib.ss_photometric_blocks(H, W)
        rec = _lib.lib.ss_photometric_record_floats()
        pairs = len(frames) * len(self.refs)
        partial = torch.empty(pairs * nb * rec, dtype=torch.float32, device=self.device)
        out = torch.empty(64, dtype=torch.float64, device=self.device)
        s = self._stream()
        history = []
        for _ in range(steps):
            self.set_calibration(cams6, rts)
            j = 0
            with torch.no_grad():
                for images in frames:
                    for slot, ref in enumerate(self.refs):
                        rel = relative_poses(rts, ref)
                        invd, (raylut, vol, _, imgs) = self._forward_slot(slot, cams6, rel, images)
                        if depth == "hard":
                            invd = 1.0 / self._hard_distance(vol)
                        cam = _cam_struct(cams6[target].tolist())
                        rt = _lib.float_array(rel[target].reshape(-1).tolist())
                        _lib.check(_lib.lib.ss_photometric_normal_equations(
                            raylut.data_ptr(), invd.data_ptr(), imgs[ref].data_ptr(), imgs[target].data_ptr(),
                            self.masks[target].data_ptr(), C.byref(cam), rt, target, int(bool(skip_selected)), H,
                            W, float(huber),
                            partial.data_ptr() + 4 * j * nb * rec, s))
                        j += 1
            _lib.check(_lib.lib.ss_gauss_newton_solve(partial.data_ptr(), j * nb, float(damping), out.data_ptr(), s))
            res = out.cpu()
            count = max(float(res[9]), 1.0)
            if res[10] != 1.0:
                break
            delta = res[:8]
            rts[target] = rts[target] @ pose_exp(delta[:6])
            cams6[target, 0] *= 1.0 + float(delta[6])
            cams6[target, 1] *= 1.0 + float(delta[7])
            history.append({"loss": float(res[8]) / count, "residuals": int(res[9]),
                            "cams6": cams6[target].clone(), "rt": rts[target].clone()})
        self.set_calibration(self.cams6, self.rts)
        return cams6, rts, history


def rotation_angle_deg(R):
    """Angle of a 3 x 3 rotation in degrees."""
    c = (torch.trace(R.double()) - 1) / 2
    return math.degrees(math.acos(max(-1.0, min(1.0, float(c)))))
