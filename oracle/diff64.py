"""FP64 differentiable restatement of the Tabbed DAS forward, for checking the
GPU adjoint kernels (test infrastructure only: imported by tests/, never by
the product).

Forward model (pycbf/cpu/pycbf.c:68-105,179-271 for nearest and cubic; the
project-defined linear kind), written with torch float64 tensors so that
torch.autograd gives reference gradients with respect to the channel data,
delays and apodisations:

  x = tautx[tx, p] + taurx[rx, p]  (summed in float32 as the kernels do, or in float64)
  w = apodtx[tx, p] * apodrx[rx, p],  used where |w| >= thresh
  j = floor((x - t0) / dt) clamped to n-1, zero outside [t0, t0 + dt (n-1)]
  a = (t0 + dt (j+1) - x) / dt,  b = (x - (t0 + dt j)) / dt
  linear: v = a y_j + b y_j+1;  cubic: v += dt^2/6 ((a^3-a) y2_j + (b^3-b) y2_j+1)
  nearest: v = y_k, k = (int)((x - t0)/dt + 0.5), zero unless -1 < (x - t0)/dt + 0.5 < n;
    with usf > 1 on the natural spline sampled at dt/usf (pycbf/cpu/pycbf.c:209-271),
    the fine positions in float32 as the upsampling kernel computes them
  korder_cubic: t = ((x - t0) - dt j) / dt, v = c0 + c1 t + c2 t^2 + c3 t^3 with the
    Hermite coefficients of pycbf/gpu/__engines__.cu:292-355 (end values y_j, y_j+1 and
    derivatives d0, d1 = rows ik0, ik0+1 of S applied to the k-sample window at iy0,
    placed as :314-328); interval n-1 is the constant y_{n-1}
  akima / makima: t = (x - (t0 + dt j)) / dt, the same Hermite form with the
    per-sample Akima derivatives (Akima's weights |m_i+1 - m_i|, |m_i-1 - m_i-2|,
    makima adding |m_i+1 + m_i| / 2, |m_i-1 + m_i-2| / 2; two extra slopes at each
    end; (m_i-1 + m_i) / 2 where both weights vanish)
with y_n = y2_n = 0 (the zero entry past the trace) and y2 the natural
spline's second derivatives (y2_0 = y2_{n-1} = 0), solved exactly in float64.

Pinned by tests/test_diff_reference.py: central finite differences of this
forward, linearity in the channel data, and agreement with the float32
oracle (oracle.das_tabbed) on the forward values.
"""
from __future__ import annotations

import torch


def _solve_141(rhs: torch.Tensor) -> torch.Tensor:
    """x with tridiag(1, 4, 1) x = rhs along the last axis (no autograd): the
    Thomas algorithm, vectorised over the leading axes (the matrix is strictly
    diagonally dominant, so no pivoting is needed)."""
    import numpy as np
    m = rhs.shape[-1]
    cp = np.empty(m)
    cp[0] = 0.25
    for i in range(1, m):
        cp[i] = 1.0 / (4.0 - cp[i - 1])
    d = torch.empty_like(rhs)
    d[..., 0] = rhs[..., 0] * 0.25
    for i in range(1, m):
        d[..., i] = (rhs[..., i] - d[..., i - 1]) * float(cp[i])
    for i in range(m - 2, -1, -1):
        d[..., i] -= float(cp[i]) * d[..., i + 1]
    return d


def natural_y2(y: torch.Tensor, dt: float) -> torch.Tensor:
    """Second derivatives of the natural cubic spline through y (..., n).
    Dense solve (differentiable); without autograd and for long traces of
    many channels, the tridiagonal sweep (_solve_141, same system)."""
    n = y.shape[-1]
    if n < 3:
        return torch.zeros_like(y)
    m = n - 2
    if not (y.requires_grad and torch.is_grad_enabled()) and y.numel() > (1 << 22):
        rhs = (y[..., 2:] - 2.0 * y[..., 1:-1] + y[..., :-2]) * (6.0 / (dt * dt))
        z = torch.zeros(y.shape[:-1] + (1,), dtype=y.dtype, device=y.device)
        return torch.cat([z, _solve_141(rhs), z], dim=-1)
    T = (torch.diag(torch.full((m,), 4.0, dtype=y.dtype, device=y.device))
         + torch.diag(torch.ones(m - 1, dtype=y.dtype, device=y.device), 1)
         + torch.diag(torch.ones(m - 1, dtype=y.dtype, device=y.device), -1))
    rhs = (y[..., 2:] - 2.0 * y[..., 1:-1] + y[..., :-2]) * (6.0 / (dt * dt))
    flat = rhs.reshape(-1, m).T
    sol = torch.linalg.solve(T, flat).T.reshape(rhs.shape)
    z = torch.zeros(y.shape[:-1] + (1,), dtype=y.dtype, device=y.device)
    return torch.cat([z, sol, z], dim=-1)


def korder_coefs(rf: torch.Tensor, S) -> torch.Tensor:
    """Hermite coefficients (..., n, 4) of the k-order cubic through rf (..., n)
    (pycbf/gpu/__engines__.cu:292-355): for interval j < n-1, c = (y_j, d0,
    3 (y_j+1 - y_j) - 2 d0 - d1, 2 (y_j - y_j+1) + d0 + d1); c = (y_{n-1}, 0,
    0, 0) for the last."""
    import numpy as np
    S = torch.as_tensor(S, dtype=rf.dtype)
    k = S.shape[0]
    n = rf.shape[-1]
    js = np.arange(n - 1)
    interior = (k - 1) // 2
    ik0 = np.where(js < k // 2, js, np.where(js > n - k // 2 - 1, k - n + js, interior))
    iy0 = np.where(js < k // 2, 0, np.where(js > n - k // 2 - 1, n - k, js - interior))
    idx = torch.as_tensor(iy0[:, None] + np.arange(k)[None, :])
    win = rf[..., idx]                                         # (..., n-1, k)
    d0 = (win * S[torch.as_tensor(ik0)]).sum(-1)
    d1 = (win * S[torch.as_tensor(ik0 + 1)]).sum(-1)
    y0, y1 = rf[..., :-1], rf[..., 1:]
    c = torch.stack([y0, d0, 3.0 * (y1 - y0) - 2.0 * d0 - d1, 2.0 * (y0 - y1) + d0 + d1], -1)
    z = torch.zeros_like(rf[..., -1:])
    last = torch.stack([rf[..., -1:], z, z, z], -1)
    return torch.cat([c, last], dim=-2)


def akima_coefs(rf: torch.Tensor, modified: bool) -> torch.Tensor:
    """Hermite coefficients (..., n, 4) of the Akima (or makima) interpolant
    through rf (..., n) on unit spacing: c = hermite(y_j, y_j+1, d_j, d_j+1),
    c = (y_{n-1}, 0, 0, 0) for the last interval."""
    m = rf[..., 1:] - rf[..., :-1]                              # (..., n-1)
    lo = torch.stack([3.0 * m[..., 0] - 2.0 * m[..., 1], 2.0 * m[..., 0] - m[..., 1]], -1)
    hi = torch.stack([2.0 * m[..., -1] - m[..., -2], 3.0 * m[..., -1] - 2.0 * m[..., -2]], -1)
    mp = torch.cat([lo, m, hi], -1)                             # slopes -2 .. n, (..., n+3)
    mm2, mm1, mp0, mp1 = mp[..., :-3], mp[..., 1:-2], mp[..., 2:-1], mp[..., 3:]
    w1 = (mp1 - mp0).abs()
    w2 = (mm1 - mm2).abs()
    if modified:
        w1 = w1 + 0.5 * (mp1 + mp0).abs()
        w2 = w2 + 0.5 * (mm1 + mm2).abs()
    ws = w1 + w2
    pos = ws > 0
    safe = torch.where(pos, ws, torch.ones_like(ws))
    d = torch.where(pos, (w1 * mm1 + w2 * mp0) / safe, 0.5 * (mm1 + mp0))   # (..., n)
    y0, y1, d0, d1 = rf[..., :-1], rf[..., 1:], d[..., :-1], d[..., 1:]
    c = torch.stack([y0, d0, 3.0 * (y1 - y0) - 2.0 * d0 - d1, 2.0 * (y0 - y1) + d0 + d1], -1)
    z = torch.zeros_like(rf[..., -1:])
    last = torch.stack([rf[..., -1:], z, z, z], -1)
    return torch.cat([c, last], dim=-2)


def upsample(rf: torch.Tensor, t0: float, dt: float, usf: int) -> torch.Tensor:
    """The natural spline through rf (..., n) at the usf (n-1) + 1 float32 fine
    positions t0 + it * (dt / usf) (zero past the last knot), in float64."""
    import numpy as np
    n = rf.shape[-1]
    ntu = usf * (n - 1) + 1
    t0f, dtf = np.float32(t0), np.float32(dt)
    dtuf = np.float32(dtf / np.float32(usf))
    xf = t0f + np.arange(ntu, dtype=np.float32) * dtuf
    ok = ~((xf < t0f) | (xf > t0f + dtf * np.float32(n - 1)))
    j = np.minimum(((xf - t0f) / dtf).astype(np.int64), n - 1)
    xd, t0d, dtd = xf.astype(np.float64), float(t0f), float(dtf)
    a = torch.as_tensor((t0d + dtd * (j + 1) - xd) / dtd)
    b = torch.as_tensor((xd - (t0d + dtd * j)) / dtd)
    zpad = torch.zeros(rf.shape[:-1] + (1,), dtype=rf.dtype)
    y = torch.cat([rf, zpad], -1)
    y2 = torch.cat([natural_y2(rf, dtd), zpad], -1)
    jt = torch.as_tensor(j)
    u = (a * y[..., jt] + b * y[..., jt + 1]
         + (dtd * dtd / 6.0) * ((a ** 3 - a) * y2[..., jt] + (b ** 3 - b) * y2[..., jt + 1]))
    return u * torch.as_tensor(ok, dtype=rf.dtype)


def das_tabbed(rf, tautx, apodtx, taurx, apodrx, t0, dt, kind="cubic", sumtype="tx_and_rx",
               thresh=0.0, x32=True, S=None, usf=1, y2=None):
    """rf (nfr, ntx, nrx, n); tables (ntx, np) / (nrx, np); returns float64
    with the sumtype layout of the GPU beamformers (frames first).  x32 sums
    the delays in float32 as the kernels do (same sample indices); x32=False
    keeps the whole function in float64 (finite-difference checks).  S: the
    (k, k) derivative operator of korder_cubic; usf: the upsampling factor
    of nearest; y2: natural_y2(rf, dt) when the caller already has it (cubic,
    repeated evaluations on the same channel data)."""
    rf = rf.to(torch.float64)
    if kind == "nearest" and usf > 1:
        # nearest on the upsampled trace at dt / usf (float32, as the kernels)
        rf = upsample(rf, t0, dt, usf)
        dt = float(torch.tensor(dt, dtype=torch.float32) / torch.tensor(float(usf), dtype=torch.float32))
    nfr, ntx, nrx, n = rf.shape
    if x32:
        # float32 delay sum, and sample indices / trace bounds from float32
        # arithmetic on the float32 time axis, as the kernels select them
        xs = tautx.to(torch.float32)[:, None, :] + taurx.to(torch.float32)[None, :, :]
        x = xs.to(torch.float64)
        t0f = torch.tensor(t0, dtype=torch.float32)
        dtf = torch.tensor(dt, dtype=torch.float32)
        q = ((xs.detach() - t0f) / dtf)
        tmaxf = t0f + dtf * torch.tensor(float(n - 1), dtype=torch.float32)
        t0, dt = float(t0f), float(dtf)
        lo_ok = xs.detach() >= t0f
        hi_ok = xs.detach() <= tmaxf
        qd = q.to(torch.float64)
    else:
        x = tautx.to(torch.float64)[:, None, :] + taurx.to(torch.float64)[None, :, :]
        qd = (x.detach() - t0) / dt
        tmax = t0 + dt * (n - 1)
        lo_ok = x.detach() >= t0
        hi_ok = x.detach() <= tmax
    w = apodtx.to(torch.float64)[:, None, :] * apodrx.to(torch.float64)[None, :, :]
    keep = (w.abs() >= thresh).to(torch.float64)
    zpad = torch.zeros(rf.shape[:-1] + (1,), dtype=rf.dtype, device=rf.device)
    y = torch.cat([rf, zpad], dim=-1)                         # (nfr, ntx, nrx, n+1)

    def take(arr, idx):
        idx = idx.clamp(0, n).expand(nfr, *idx.shape)
        return torch.gather(arr, -1, idx)

    if kind == "nearest":
        # (int)(q + 0.5) with the 0.5 added in double: q in (-1.5, -0.5) selects
        # sample 0 (pycbf/cpu/pycbf.c:264,267)
        d = qd + 0.5
        inside = ((d > -1.0) & (d < n)).to(torch.float64)
        v = take(y, torch.trunc(d).clamp(0, n - 1).long())
    elif kind in ("korder_cubic", "akima", "makima"):
        inside = (lo_ok & hi_ok).to(torch.float64)
        j = torch.floor(qd).clamp(0, n - 1).long()
        if kind == "korder_cubic":
            t = ((x - t0) - dt * j.to(torch.float64)) / dt
            C = korder_coefs(rf, S)                            # (nfr, ntx, nrx, n, 4)
        else:
            t = (x - (t0 + dt * j.to(torch.float64))) / dt
            C = akima_coefs(rf, kind == "makima")
        cj = torch.gather(C, 3, j[None, ..., None].expand(nfr, *j.shape, 4))
        v = cj[..., 0] + t * (cj[..., 1] + t * (cj[..., 2] + t * cj[..., 3]))
    else:
        inside = (lo_ok & hi_ok).to(torch.float64)
        j = torch.floor(qd).clamp(0, n - 1).long()
        jf = j.to(torch.float64)
        a = (t0 + dt * (jf + 1) - x) / dt
        b = (x - (t0 + dt * jf)) / dt
        v = a * take(y, j) + b * take(y, j + 1)
        if kind == "cubic":
            y2 = torch.cat([natural_y2(rf, dt) if y2 is None else y2.to(rf.dtype), zpad],
                           dim=-1)
            v = v + (dt * dt / 6.0) * ((a ** 3 - a) * take(y2, j) + (b ** 3 - b) * take(y2, j + 1))
    s = v * (w * keep * inside)[None]                        # (nfr, ntx, nrx, np)
    if sumtype == "none":
        return s
    if sumtype == "tx_only":
        return s.sum(dim=1)
    if sumtype == "rx_only":
        return s.sum(dim=2)
    return s.sum(dim=(1, 2))


def ray_delays(slow, geo, steer, xorig, xele, pvec, c0, nsteps, chunk=4096):
    """Straight-ray delays through a bilinear, edge-extended slowness map, in
    float64 (restating pycbf_ray_delays): returns (tautx (ntx, np), taurx
    (nrx, np)); differentiable with respect to ``slow`` (nzm, nxm).
    ``geo`` has nzm, nxm, xm0, zm0, dxm, dzm."""
    slow = slow.to(torch.float64)
    dev = slow.device
    steer = torch.as_tensor(steer, dtype=torch.float64, device=dev)
    xorig = torch.as_tensor(xorig, dtype=torch.float64, device=dev)
    xele = torch.as_tensor(xele, dtype=torch.float64, device=dev)
    pvec = torch.as_tensor(pvec, dtype=torch.float64, device=dev)
    nzm, nxm = geo.nzm, geo.nxm
    t = (torch.arange(nsteps, dtype=torch.float64, device=dev) + 0.5) / nsteps

    def sample(x, z):
        u = ((x - geo.xm0) / geo.dxm - 0.5).clamp(0.0, nxm - 1.0)
        v = ((z - geo.zm0) / geo.dzm - 0.5).clamp(0.0, nzm - 1.0)
        k0 = torch.floor(u).clamp(max=max(nxm - 2, 0)).long()
        i0 = torch.floor(v).clamp(max=max(nzm - 2, 0)).long()
        wx = (u - k0) if nxm > 1 else torch.zeros_like(u)
        wz = (v - i0) if nzm > 1 else torch.zeros_like(v)
        k1 = (k0 + 1).clamp(max=nxm - 1)
        i1 = (i0 + 1).clamp(max=nzm - 1)
        f = slow.reshape(-1)
        s00, s01 = f[i0 * nxm + k0], f[i0 * nxm + k1]
        s10, s11 = f[i1 * nxm + k0], f[i1 * nxm + k1]
        top = s00 + wx * (s01 - s00)
        bot = s10 + wx * (s11 - s10)
        return top + wz * (bot - top)

    out_tx, out_rx = [], []
    for p0 in range(0, pvec.shape[0], chunk):
        px = pvec[p0:p0 + chunk, 0][None, :]
        pz = pvec[p0:p0 + chunk, 1][None, :]
        sn, cs = torch.sin(steer)[:, None], torch.cos(steer)[:, None]
        qx = px - pz * (sn / cs)
        L = pz / cs
        fire = (qx - xorig[:, None]) * sn / c0
        xs = qx[..., None] + t * (px - qx)[..., None]
        zs = (t * pz[..., None]).expand_as(xs)
        out_tx.append(fire + L * sample(xs, zs).mean(dim=-1))
        qx = xele[:, None].expand(-1, px.shape[1])
        L = torch.sqrt((px - qx) ** 2 + pz ** 2)
        xs = qx[..., None] + t * (px - qx)[..., None]
        zs = (t * pz[..., None]).expand_as(xs)
        out_rx.append(L * sample(xs, zs).mean(dim=-1))
    return torch.cat(out_tx, dim=1), torch.cat(out_rx, dim=1)
