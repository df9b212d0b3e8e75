"""PyTorch float64 autograd reference of the differentiation form.

TEST INFRASTRUCTURE ONLY.  The forward restates, with differentiable torch
ops, the path the GPU kernels differentiate: unit rays (utils.py:57-73),
sweep + clamped cost with torch's own grid_sample (depth_estimation.py:
141-175), the ISB cost filter (isb_filter.py:72-136, isb_filter.cu:44-330)
with its guide pyramid taken from the fp32 oracle (the bytes the GPU forward
produces) and its stencil weights tabulated from those guides, then the
soft-argmin readout.  torch.autograd on it is the reference gradient the
hand-written GPU adjoints are checked against.
"""
import numpy as np
import torch
import torch.nn.functional as F

f64 = torch.float64


def _sqrt_safe(t):
    """sqrt with a zero derivative where the argument is 0 (the clamp of
    utils.py:66 acts there)."""
    pos = t > 0
    return torch.where(pos, torch.sqrt(torch.where(pos, t, torch.ones_like(t))), torch.zeros_like(t))


def unproject(cam, H, W):
    fs0, fs1, cs0, cs1, xi, al = cam.unbind()
    v, u = torch.meshgrid(torch.arange(H, dtype=f64), torch.arange(W, dtype=f64), indexing="ij")
    mx = (u - cs0) / fs0
    my = (v - cs1) / fs1
    r2 = mx * mx + my * my
    t = torch.clamp(1 - (2 * al - 1) * r2, min=0)
    mz = (1 - al * al * r2) / (al * _sqrt_safe(t) + 1 - al)
    k = (mz * xi + torch.sqrt(mz * mz + (1 - xi * xi) * r2)) / (mz * mz + r2)
    return torch.stack([k * mx, k * my, k * mz - xi], -1)


def project(cam, q):
    fs0, fs1, cs0, cs1, xi, al = cam.unbind()
    x, y, z = q.unbind(-1)
    xy2 = x * x + y * y
    d1 = torch.sqrt(xy2 + z * z)
    c = xi * d1 + z
    d2 = torch.sqrt(xy2 + c * c)
    nrm = al * d2 + (1 - al) * c
    return fs0 * x / nrm + cs0, fs1 * y / nrm + cs1


def sweep_cost(images, cams, rel, ref, sel, cands):
    """Clamped cost [N, H, W]: images C x [H, W, 3], cams [C, 6], rel [C, 3, 4]
    (camera k <- reference), sel [H, W] (-1 = none), cands [N]."""
    H, W, _ = images[0].shape
    N = cands.shape[0]
    ray = unproject(cams[ref], H, W)
    P = cands[:, None, None, None] * ray[None]
    sample = torch.zeros((N, H, W, 3), dtype=f64)
    for k in range(len(images)):
        m = sel == k
        if not bool(m.any()):
            continue
        q = P @ rel[k][:, :3].T + rel[k][:, 3]
        u, v = project(cams[k], q)
        g = torch.stack([((u + 0.5) / W) * 2 - 1, ((v + 0.5) / H) * 2 - 1], -1).reshape(1, N * H, W, 2)
        s = F.grid_sample(images[k].permute(2, 0, 1)[None], g, mode="bilinear", padding_mode="zeros",
                          align_corners=False)
        s = s[0].reshape(3, N, H, W).permute(1, 2, 3, 0)
        sample = torch.where(m[None, :, :, None], s, sample)
    cost = (sample - images[ref][None]).abs().sum(-1)
    return torch.clamp(cost, max=500)


# ---- ISB filter with tabulated weights --------------------------------------------
_T1 = [(0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (1, 2), (2, 0), (2, 1), (2, 2)]
_T2 = [(1, 1), (2, 1), (1, 0), (2, 0), (1, 2), (2, 2)]
_T3 = [(1, 1), (1, 2), (0, 1), (0, 2), (2, 1), (2, 2)]
_T4 = [(1, 1), (1, 2), (2, 1), (2, 2)]
_B1 = [1, 1, 1, 1, 8, 1, 1, 1, 1]
_B23 = [8, 8, 1, 1, 1, 1]
_B4 = [1, 1, 1, 1]


def _sad(a, b):
    return np.abs(a.astype(np.int64) - b.astype(np.int64)).sum(-1)


def filter_tables(guide0, sigma_i, sigma_s, oracle32):
    """Per level: down-pass gathers (centre, 8 neighbours, weights, 1/ws) and
    up-pass gathers (per fine pixel: up to 9 coarse taps with normalised
    weights, and the weight of its own down-pass value)."""
    H, W, _ = guide0.shape
    L = oracle32.levels(W, H)
    vi = 1 / (2 * sigma_i * sigma_i)
    vs = 1 / (2 * sigma_s * sigma_s)
    sizes = [(H, W)] + [(int(np.ceil(H / 2 ** s)), int(np.ceil(W / 2 ** s))) for s in range(1, L)]
    G = [guide0]
    down = [None]
    for s in range(1, L):
        hin, win = sizes[s - 1]
        hout, wout = sizes[s]
        g, _ = oracle32.isb_down(G[s - 1], np.zeros((0, hin, win)), hout, wout, vi)
        G.append(g)
        y, x = np.meshgrid(np.arange(hout), np.arange(wout), indexing="ij")
        x2, y2 = 2 * x, 2 * y
        xp, yp = np.minimum(x2 + 1, win - 1), np.minimum(y2 + 1, hin - 1)
        xm, ym = np.maximum(x2 - 1, 0), np.maximum(y2 - 1, 0)
        nbs = [(ym, xm), (y2, xm), (yp, xm), (yp, x2), (yp, xp), (y2, xp), (ym, xp), (ym, x2)]
        gp = G[s - 1]
        gc = gp[y2, x2]
        w = []
        for j, (r, c) in enumerate(nbs):
            d = _sad(gc, gp[r, c]).astype(np.float64)
            w.append((0.5 if j & 1 else 0.25) * (np.exp(-d * d * vi) + 0.01))
        w = np.stack(w, -1)
        ws = 1 + w.sum(-1)
        down.append(dict(ctr=torch.from_numpy((y2 * win + x2).ravel()),
                         nb=torch.from_numpy(np.stack([r * win + c for r, c in nbs], -1).reshape(-1, 8)),
                         w=torch.from_numpy(w.reshape(-1, 8)), inv=torch.from_numpy(1 / ws.ravel())))
    up = [None] * L
    Gu = G[L - 1]
    for s in range(L - 2, -1, -1):
        hin, win = sizes[s + 1]
        hio, wio = sizes[s]
        dist = 2.0 ** s - 0.5
        wdn = np.exp(-(dist * dist) * vs)
        wup = 1 - wdn
        taps = np.zeros((hio * wio, 9), np.int64)
        tw = np.zeros((hio * wio, 9))
        selfw = np.zeros(hio * wio)
        for y in range(hin):
            rows = [max(y - 1, 0), y, min(y + 1, hin - 1)]
            for x in range(win):
                cols = [max(x - 1, 0), x, min(x + 1, win - 1)]
                x2, y2 = 2 * x, 2 * y
                xpu, ypu = min(x2 + 1, wio - 1), min(y2 + 1, hio - 1)
                fine = [(y2, x2), (ypu, x2), (y2, xpu), (ypu, xpu)]
                doit = [True, y2 != ypu, x2 != xpu, y2 != ypu and x2 != xpu]
                for t, (tab, base) in enumerate([(_T1, _B1), (_T2, _B23), (_T3, _B23), (_T4, _B4)]):
                    if not doit[t]:
                        continue
                    fy, fx = fine[t]
                    gf = G[s][fy, fx]
                    ws = wup
                    ww, ii = [], []
                    for (r, c), b in zip(tab, base):
                        gn = Gu[rows[r], cols[c]]
                        d = float(_sad(gn, gf))
                        wk = b * (wdn * (np.exp(-d * d * vi) + 0.01))
                        ww.append(wk)
                        ii.append(rows[r] * win + cols[c])
                        ws += wk
                    f = fy * wio + fx
                    taps[f, :len(ii)] = ii
                    tw[f, :len(ww)] = np.array(ww) / ws
                    selfw[f] = wup / ws
        up[s] = dict(taps=torch.from_numpy(taps), w=torch.from_numpy(tw), self=torch.from_numpy(selfw))
        Gu, _ = oracle32.isb_up(Gu, np.zeros((0, hin, win)), G[s], np.zeros((0, hio, wio)), wup, wdn, vi)
    return dict(L=L, sizes=sizes, down=down, up=up)


def isb_filter(cost, tab):
    """cost [N, H, W] float64 -> filtered [N, H, W] (differentiable)."""
    N, H, W = cost.shape
    L = tab["L"]
    Cs = [cost.reshape(N, -1)]
    for s in range(1, L):
        d = tab["down"][s]
        prev = Cs[s - 1]
        Cs.append((prev[:, d["ctr"]] + (prev[:, d["nb"]] * d["w"]).sum(-1)) * d["inv"])
    Fl = Cs[L - 1]
    for s in range(L - 2, -1, -1):
        u = tab["up"][s]
        Fl = u["self"] * Cs[s] + (Fl[:, u["taps"]] * u["w"]).sum(-1)
    return Fl.reshape(N, H, W)


def soft_argmin(Fv, beta, inv0, inv1):
    N = Fv.shape[0]
    sm = torch.softmax(-beta * Fv, dim=0)
    iota = (torch.arange(N, dtype=f64)[:, None, None] * sm).sum(0)
    return inv0 + (inv1 - inv0) * iota / (N - 1)


def forward(images, cams, rel, ref, sel, cands, tab, beta, inv0, inv1):
    """Inverse distance [H, W] of reference `ref`."""
    return soft_argmin(isb_filter(sweep_cost(images, cams, rel, ref, sel, cands), tab), beta, inv0, inv1)
