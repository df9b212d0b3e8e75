"""Seeded synthetic rigs and scenes: the data source of the tests and benchmark.

A rig is ``n`` double-sphere fisheye cameras on the corners of a square of side
``baseline`` (metres), optical axes yawed 45 + k*90 degrees outward.  The scene
is a closed star-shaped surface r(dir) = exp(3-octave value noise) clipped to
[0.8, 20] m plus axis-aligned boxes at 1.2-6 m, all carrying 4-octave value-noise
colour.  Images are ray-cast at the original resolution (uint8, 3 channels)
with a circular validity mask of the given field of view.  The rig is also
available as a basalt calibration JSON so that ``parse_json_calib`` applies.

Camera frame: x right, y down, z forward; poses map camera to rig coordinates.
"""
import math

import torch
import torch.nn.functional as F


def _hash01(ix, iy, iz, seed):
    h = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761)
    h = h & 0xFFFFFFFF
    h = ((h ^ (h >> 16)) * 0x45D9F3B) & 0xFFFFFFFF
    h = ((h ^ (h >> 16)) * 0x45D9F3B) & 0xFFFFFFFF
    h = h ^ (h >> 16)
    return (h & 0xFFFFFF).to(torch.float32) / float(0xFFFFFF)


def value_noise(p, seed):
    """Trilinear value noise in [0, 1] on the integer lattice (smoothstep)."""
    f = torch.floor(p)
    t = p - f
    s = t * t * (3 - 2 * t)
    i = f.to(torch.int64)
    out = torch.zeros(p.shape[:-1], dtype=p.dtype, device=p.device)
    for dx in (0, 1):
        wx = s[..., 0] if dx else 1 - s[..., 0]
        for dy in (0, 1):
            wy = s[..., 1] if dy else 1 - s[..., 1]
            for dz in (0, 1):
                wz = s[..., 2] if dz else 1 - s[..., 2]
                out = out + wx * wy * wz * _hash01(i[..., 0] + dx, i[..., 1] + dy, i[..., 2] + dz, seed)
    return out


def fbm(p, octaves, seed):
    acc, norm, amp, freq = 0.0, 0.0, 1.0, 1.0
    for o in range(octaves):
        acc = acc + amp * value_noise(p * freq, seed + 17 * o)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return acc / norm


def _rot_y(psi):
    c, s = math.cos(psi), math.sin(psi)
    return [[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]]


class SyntheticRig:
    def __init__(self, seed=0, resolution=1216, baseline=0.26, n_cameras=4, fx_ratio=0.30, xi=-0.18, alpha=0.59,
                 fov_deg=220.0, n_boxes=8):
        self.seed = int(seed)
        self.resolution = int(resolution)
        self.baseline = float(baseline)
        self.xi, self.alpha = float(xi), float(alpha)
        self.fov = math.radians(fov_deg)
        self.f = fx_ratio * resolution
        self.c = resolution / 2.0
        self.poses = []
        for k in range(n_cameras):
            psi = math.radians(45.0 + 90.0 * k)
            rad = self.baseline / math.sqrt(2.0)
            self.poses.append((psi, [rad * math.sin(psi), 0.0, rad * math.cos(psi)]))
        g = torch.Generator().manual_seed(self.seed)
        self.boxes = []
        for _ in range(n_boxes):
            d = torch.randn(3, generator=g, dtype=torch.float64)
            d[1] *= 0.4  # keep most boxes near the horizon
            d = d / d.norm()
            dist = 1.2 + 4.8 * torch.rand(1, generator=g, dtype=torch.float64).item()
            half = 0.1 + 0.3 * torch.rand(3, generator=g, dtype=torch.float64)
            ctr = d * dist
            self.boxes.append(((ctr - half).tolist(), (ctr + half).tolist(), 1000 + len(self.boxes)))

    # -- calibration ----------------------------------------------------------
    def basalt_json(self):
        """{'value0': ...} basalt calibration of the rig."""
        ext, intr = [], []
        for psi, t in self.poses:
            ext.append({"px": t[0], "py": t[1], "pz": t[2], "qx": 0.0, "qy": math.sin(psi / 2), "qz": 0.0,
                        "qw": math.cos(psi / 2)})
            intr.append({"camera_type": "ds", "intrinsics": {"fx": self.f, "fy": self.f, "cx": self.c, "cy": self.c,
                                                             "xi": self.xi, "alpha": self.alpha}})
        return {"value0": {"T_imu_cam": ext, "intrinsics": intr,
                           "resolution": [[self.resolution, self.resolution]] * len(self.poses)}}

    # -- geometry ---------------------------------------------------------------
    def _unproject(self, u, v):
        """Double-sphere unprojection in the working dtype; returns unit rays
        and the field-of-view mask."""
        mx, my = (u - self.c) / self.f, (v - self.c) / self.f
        r2 = mx * mx + my * my
        a, xi = self.alpha, self.xi
        arg = 1 - (2 * a - 1) * r2
        ok = arg >= 0
        mz = (1 - a * a * r2) / (a * torch.sqrt(torch.clamp(arg, min=0)) + 1 - a)
        k = (mz * xi + torch.sqrt(torch.clamp(mz * mz + (1 - xi * xi) * r2, min=0))) / (mz * mz + r2)
        p = torch.stack([k * mx, k * my, k * mz - xi], dim=-1)
        p = p / p.norm(dim=-1, keepdim=True)
        ok = ok & (torch.acos(torch.clamp(p[..., 2], -1, 1)) <= self.fov / 2)
        return p, ok

    def _surface_radius(self, d):
        n = fbm(d * 2.0, 3, self.seed)
        return torch.clamp(torch.exp(0.4 + 3.0 * n), 0.8, 20.0)

    def raycast(self, origin, dirs):
        """Distance along unit ``dirs`` from ``origin`` (3,) to the scene and
        the hit points."""
        o = torch.as_tensor(origin, dtype=dirs.dtype, device=dirs.device)
        t = self._surface_radius(dirs)
        for _ in range(12):
            p = o + t[..., None] * dirs
            rho = p.norm(dim=-1)
            t = torch.clamp(t + (self._surface_radius(p / rho[..., None]) - rho), min=0.05)
        best = t
        tex = torch.full_like(t, float(self.seed + 100))
        inv = 1.0 / torch.where(dirs.abs() < 1e-12, torch.full_like(dirs, 1e-12), dirs)
        for lo, hi, tseed in self.boxes:
            lo_t = torch.tensor(lo, dtype=dirs.dtype, device=dirs.device)
            hi_t = torch.tensor(hi, dtype=dirs.dtype, device=dirs.device)
            t1, t2 = (lo_t - o) * inv, (hi_t - o) * inv
            tn = torch.minimum(t1, t2).max(dim=-1).values
            tf = torch.maximum(t1, t2).min(dim=-1).values
            hit = (tn <= tf) & (tn > 0) & (tn < best)
            best = torch.where(hit, tn, best)
            tex = torch.where(hit, torch.full_like(tex, float(tseed)), tex)
        return best, o + best[..., None] * dirs, tex

    def _colour(self, pts, tex):
        out = []
        for ch in range(3):
            c = torch.zeros(pts.shape[:-1], dtype=pts.dtype, device=pts.device)
            for tseed in torch.unique(tex).tolist():
                m = tex == tseed
                val = fbm(pts * 3.0, 4, int(tseed) * 7 + 31 * ch)
                c = torch.where(m, val, c)
            out.append(torch.clamp(0.5 + 1.8 * (c - 0.5), 0, 1) * 255)
        return torch.stack(out, dim=-1)

    def render(self, device="cpu", dtype=torch.float32, rows_per_chunk=256):
        """Ray-cast every camera: (images uint8 [C, R, R, 3], masks uint8 [C, R, R])."""
        R = self.resolution
        imgs = torch.zeros([len(self.poses), R, R, 3], dtype=torch.uint8, device=device)
        masks = torch.zeros([len(self.poses), R, R], dtype=torch.uint8, device=device)
        for k, (psi, t) in enumerate(self.poses):
            rot = torch.tensor(_rot_y(psi), dtype=dtype, device=device)
            for r0 in range(0, R, rows_per_chunk):
                r1 = min(R, r0 + rows_per_chunk)
                v, u = torch.meshgrid(torch.arange(r0, r1, device=device, dtype=dtype),
                                      torch.arange(R, device=device, dtype=dtype), indexing="ij")
                p, ok = self._unproject(u, v)
                dirs = p @ rot.T
                _, pts, tex = self.raycast(t, dirs)
                col = self._colour(pts, tex)
                col = torch.where(ok[..., None], col, torch.zeros_like(col))
                imgs[k, r0:r1] = col.to(torch.uint8)
                masks[k, r0:r1] = ok.to(torch.uint8) * 255
        return imgs, masks

    def panorama_inverse_distance(self, viewpoint, cols, rows, device="cpu", dtype=torch.float32):
        """Ground-truth inverse distance on the equirectangular grid of the
        stitcher (stitcher.cu:346-354) seen from ``viewpoint``."""
        y, x = torch.meshgrid(torch.arange(rows, device=device, dtype=torch.float64),
                              torch.arange(cols, device=device, dtype=torch.float64), indexing="ij")
        phi = (y + 0.5) * math.pi / rows - math.pi / 2
        theta = (x + 0.5) * 2 * math.pi / cols + math.pi
        d = torch.stack([torch.cos(phi) * torch.sin(theta), torch.sin(phi), torch.cos(phi) * torch.cos(theta)], -1)
        dist, _, _ = self.raycast(list(map(float, viewpoint)), d.to(dtype))
        return 1.0 / dist


def resize_bilinear(img, cols, rows):
    """Bilinear resize with half-pixel centres and clamped borders (the
    INTER_LINEAR behaviour of the reference's cv2.resize calls,
    utils.py:198-201), float32 in and out, HWC or HW."""
    x = img.to(torch.float32)
    hw = x.dim() == 2
    x = x[None, None] if hw else x.permute(2, 0, 1)[None]
    y = F.interpolate(x, size=(rows, cols), mode="bilinear", align_corners=False, antialias=False)
    return y[0, 0] if hw else y[0].permute(1, 2, 0).contiguous()


def make_frame(rig, matching_resolution, rgb_to_stitch_resolution, references, device="cpu", rendered=None):
    """Rendered rig -> (images_to_match, images_to_stitch, masks) as the
    reference's main.py:72-81 and read_input_images (utils.py:157-204) hand
    them to the estimator: float32 HWC in [0, 255] and [1, rows, cols] masks."""
    imgs, masks = rendered if rendered is not None else rig.render(device)
    mc, mr = matching_resolution
    sc, sr = rgb_to_stitch_resolution
    to_match = [resize_bilinear(imgs[k], mc, mr).to(device) for k in range(imgs.shape[0])]
    to_stitch = [resize_bilinear(imgs[k], sc, sr).to(device) for k in references]
    mks = [(resize_bilinear(masks[k], mc, mr) / 255).unsqueeze(0).to(device) for k in range(masks.shape[0])]
    return to_match, to_stitch, mks


# ---------------------------------------------------------------------------
# Named workloads (sizes as the reference passes them: (cols, rows))
# ---------------------------------------------------------------------------
WORKLOADS = {
    # 512^2 matching, 1024 x 512 panorama (the CPU-sized configuration)
    "c0": dict(matching=(512, 512), stitch=(1216, 1216), pano=(1024, 512), N=32, refs=(0, 2), sigma_i=10.0,
               sigma_s=12.5, resolution=1216, seed=0),
    # main.py's defaults: 1024^2 matching, 2048 x 1024 panorama
    "c1": dict(matching=(1024, 1024), stitch=(1216, 1216), pano=(2048, 1024), N=32, refs=(0, 2), sigma_i=10.0,
               sigma_s=25.0, resolution=1216, seed=0),
    # 64 candidates, every camera a reference
    "c2": dict(matching=(1024, 1024), stitch=(1216, 1216), pano=(2048, 1024), N=64, refs=(0, 1, 2, 3),
               sigma_i=10.0, sigma_s=25.0, resolution=1216, seed=0),
}


def _frame_of(rig, rendered, matching, stitch, pano, N, refs, sigma_i, sigma_s, device):
    from .utils import parse_json_calib

    calibs = parse_json_calib(rig.basalt_json()["value0"], list(matching), device)
    viewpoint = torch.zeros([3], device=device)
    for r in refs:
        viewpoint += calibs[r].rt[:3, 3] / len(refs)
    to_match, to_stitch, masks = make_frame(rig, matching, stitch, list(refs), device=device, rendered=rendered)
    kw = dict(min_dist=0.55, max_dist=100, candidate_count=N, references_indices=list(refs),
              reprojection_viewpoint=viewpoint, masks=masks, matching_resolution=tuple(matching),
              rgb_to_stitch_resolution=tuple(stitch), panorama_resolution=tuple(pano), sigma_i=sigma_i,
              sigma_s=sigma_s)
    return calibs, kw, [t.contiguous() for t in to_match], [t.contiguous() for t in to_stitch]


def workload_frame(wl, device, rendered=None):
    """One frame of a workload (a WORKLOADS entry or a dict like one): the
    seeded rig ray-cast on ``device`` (or ``rendered`` = (images, masks)
    from an earlier render), its calibrations parsed at the matching
    resolution, the viewpoint at the references' centre (main.py:67-70).
    Returns (calibrations, RGBD_Estimator keyword arguments without
    ``device``, images_to_match, images_to_stitch, rendered)."""
    rig = SyntheticRig(seed=wl["seed"], resolution=wl["resolution"], baseline=wl.get("baseline", 0.26))
    if rendered is None:
        rendered = rig.render(device)
    calibs, kw, tm, ts = _frame_of(rig, rendered, wl["matching"], wl["stitch"], wl["pano"], wl["N"], wl["refs"],
                                   wl["sigma_i"], wl["sigma_s"], device)
    return calibs, kw, tm, ts, rendered


def ragged_rigs(device, count=16, seed=3):
    """The ragged batching workload: ``count`` rigs, each a seeded scene
    (seeds 100, 101, ...) with its own baseline drawn from U(0.15, 0.40) m,
    matching resolution from {512, 768, 1024}^2, candidate count from {16,
    32, 48, 64} and 2 to 4 reference cameras; 1216^2 stitch input, 1024 x 512
    panorama, sigma_i 10, sigma_s 12.5 x (resolution / 512).  Returns one
    (calibrations, keyword arguments, images_to_match, images_to_stitch,
    (resolution, N, references)) per rig."""
    import random

    rng = random.Random(seed)
    out = []
    for i in range(count):
        res = rng.choice([512, 768, 1024])
        N = rng.choice([16, 32, 48, 64])
        refs = {2: [0, 2], 3: [0, 1, 2], 4: [0, 1, 2, 3]}[rng.choice([2, 3, 4])]
        rig = SyntheticRig(seed=100 + i, resolution=1216, baseline=rng.uniform(0.15, 0.40))
        rendered = rig.render(device)
        calibs, kw, tm, ts = _frame_of(rig, rendered, (res, res), (1216, 1216), (1024, 512), N, refs, 10.0,
                                       12.5 * res / 512, device)
        out.append((calibs, kw, tm, ts, (res, N, len(refs))))
    return out
