#!/usr/bin/env python3
"""Throughput of the beamforming hot path on one MI355X.

Workload (one step = one pass of the hot path over one batch): Synthetic-point
DAS with natural-cubic interpolation, 75 plane waves over +-16 deg encoded as
synthetic points at -10 m, 128 receive elements (0.3 mm pitch, f/1.5
acceptance), 4096 float32 samples per channel at 40 MHz, a 512 x 512 pixel grid
over z 2-50 mm and x +-19 mm, a batch of 16 independently seeded frames (50
point scatterers + 1e4 speckle scatterers each, 30 dB SNR).  RF and geometry
are resident in HBM before the timed region; each step runs the spline
prefilter, the delay tables and the gather kernel and writes the 16 images.

Prints one JSON line (metric / value in Gsamples/s of tx*rx*px*frames).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Gsamples/s (tx·rx·px) cubic DAS + SLSC px/s on 1 MI355X vs plain GPU & CPU"
HBM_PEAK_GBPS = 8000.0
BF16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak
# float32 wave64 VALU instructions: one per CU per cycle (a 16-lane SIMD takes
# 4 cycles per instruction, 4 SIMDs per CU), 256 CUs at 2.4 GHz
VALU_PEAK_GINSTR = 256 * 2.4

NTX, NRX, NT, NX, NZ = 75, 128, 4096, 512, 512


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload_geometry():
    from pycbf_amd import sim
    xele = sim.linear_array(NRX)
    steers = np.radians(np.linspace(-16, 16, NTX))
    _, _, pnts = sim.grid(NX, NZ, (-19e-3, 19e-3), (2e-3, 50e-3))
    params = sim.pw_synthetic(steers, xele, pnts)
    return xele, steers, pnts, params


def make_rf(steers, xele, nframes, device, seed=1234):
    import torch
    from pycbf_amd import sim
    rf = torch.empty((nframes, NTX, NRX, NT), dtype=torch.float32, device=device)
    for f in range(nframes):
        xs, zs, amp = sim.scatterers(50, 10_000, (-19e-3, 19e-3), (2e-3, 50e-3), seed + f)
        rf[f] = sim.simulate_rf_torch(steers, xele, NT, xs, zs, amp, device, seed=seed + f)
        log(f"simulated frame {f + 1}/{nframes}")
    torch.cuda.synchronize()
    return rf


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles",
                               "das_gather_synthetic_cubic_75x128x512x512_nt4096_16fr.md")


def profile_figures(frames):
    """Per-launch figures of the gather from the committed rocprofv3 summary,
    when it was measured for this frame count: {"traffic": HBM-side bytes
    (FETCH_SIZE + WRITE_SIZE scaled by a copy kernel of known size),
    "valu": wave-level VALU instructions (SQ_INSTS_VALU)}; empty otherwise."""
    if not os.path.exists(TRAFFIC_PROFILE):
        return {}
    vals = {}
    keys = ("traffic_bytes_per_launch", "valu_instructions_per_launch", "frames")
    with open(TRAFFIC_PROFILE) as fh:
        for line in fh:
            if ":" in line and line.split(":")[0] in keys:
                k, v = line.split(":", 1)
                vals[k] = float(v)
    if vals.get("frames") != frames:
        return {}
    return {"traffic": vals.get("traffic_bytes_per_launch"),
            "valu": vals.get("valu_instructions_per_launch")}


def cpu_info():
    """Host CPU of this run: model name, logical CPUs, CPUs this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "cpu_count": os.cpu_count(), "usable": usable}


def cpu_baseline(frames, params, stride, threads):
    """Reference C cubic DAS (oracle/_ref, pycbf/cpu/pycbf.c) -- or the oracle
    restatement when the reference build is absent -- on the given host frames
    over a strided pixel subset, channels split over `threads` host threads
    with private accumulators (the reference class's worker pool,
    pycbf/cpu/__core_bf_classes__.py:99-114).  Returns (Gsamples/s, threads,
    kind, sample description, per-frame float64 images)."""
    import ctypes as ct

    import oracle
    pnts = params["pnts"][::stride].copy()
    p = params
    tabs = oracle.synthetic_tables(p["ovectx"], p["nvectx"], p["t0tx"], p["alatx"], p["doftx"],
                                   p["ovecrx"], p["nvecrx"], p["alarx"], p["c0"], pnts)
    tautx, apodtx, taurx, apodrx = tabs
    npx = pnts.shape[0]
    R = oracle.ref()
    kind = "reference" if R is not None else "port"
    fn = R.beamform_cubic if R is not None else oracle.lib().orc_bf_cubic
    pf = ct.POINTER(ct.c_float)
    dt = float(np.float32(1.0 / 40e6))
    chans = [(t, r) for t in range(NTX) for r in range(NRX)]

    def work(o, k, rf0):
        op = o.ctypes.data_as(pf)
        for idx in range(k, len(chans), threads):
            t, r = chans[idx]
            fn(0.0, dt, NT, rf0[t, r].ctypes.data_as(pf), npx, 1e-2,
               tautx[t].ctypes.data_as(pf), apodtx[t].ctypes.data_as(pf),
               taurx[r].ctypes.data_as(pf), apodrx[r].ctypes.data_as(pf), op)

    wall = 0.0
    images = []
    for rf0 in frames:
        rf0 = np.ascontiguousarray(rf0, np.float32)
        outs = [np.zeros(npx, np.float32) for _ in range(threads)]
        ths = [threading.Thread(target=work, args=(outs[k], k, rf0)) for k in range(threads)]
        t0 = time.perf_counter()
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        wall += time.perf_counter() - t0
        images.append(np.sum(np.stack(outs).astype(np.float64), axis=0))
    samples = NTX * NRX * npx * len(frames)
    desc = (f"{len(frames)} frames, {NTX}x{NRX} channels, every {stride}th of {NX * NZ} pixels "
            f"({npx} px, {samples / 1e9:.2f} G samples), cubic, {wall:.2f} s wall")
    return samples / wall / 1e9, threads, kind, desc, images


def count_contributing(bf, thresh):
    """(tx, rx, px) samples of one frame that contribute: weight passes the
    threshold and the delay lies on the trace (the others cost the staged
    gather no table reads)."""
    import torch
    ttx, atx, trx, arx = bf.tables()
    tmax = float(np.float32(0.0) + np.float32(1.0 / 40e6) * np.float32(NT - 1))
    n = 0
    for t in range(ttx.shape[0]):
        w = (atx[t][None, :] * arx).abs() >= thresh
        x = ttx[t][None, :] + trx
        n += int((w & (x >= 0.0) & (x <= tmax)).sum())
    del ttx, atx, trx, arx
    torch.cuda.empty_cache()
    return n


def time_variant(env, value, params, rf, ref_out, steps):
    """Times the same 16-frame step with another gather selected by an
    environment switch of the library; returns (ms per step, max error
    relative to the headline images' peak)."""
    import torch

    from pycbf_amd.gpu import SyntheticBeamformer
    os.environ[env] = value
    try:
        bfv = SyntheticBeamformer(**params, t0=0.0, dt=1.0 / 40e6, nt=NT,
                                  interp={"kind": "cubic"}, sumtype="tx_and_rx",
                                  device=rf.device)
        outv = torch.empty_like(ref_out)
        bfv(rf, out_as_numpy=False, buffer=outv)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            bfv(rf, out_as_numpy=False, buffer=outv)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
    finally:
        del os.environ[env]
    err = float((outv - ref_out).abs().max() / ref_out.abs().max())
    del bfv, outv
    torch.cuda.empty_cache()
    return ms, err


def time_tabbed(bf, rf, ref_out, steps):
    """Times the same 16-frame step through a TabbedBeamformer on the
    geometry's own delay / apodisation tables (pycbf's table-driven class);
    returns (ms per step, max error relative to the headline images' peak)."""
    import torch

    from pycbf_amd.gpu import TabbedBeamformer
    from pycbf_amd.gpu import _engine as E
    tt, at, tr, ar = bf.tables()
    bft = TabbedBeamformer(tautx=tt, apodtx=at, taurx=tr, apodrx=ar, t0=0.0, dt=1.0 / 40e6,
                           nt=NT, interp={"kind": "cubic"}, sumtype="tx_and_rx",
                           device=rf.device)
    del tt, at, tr, ar
    outv = torch.empty_like(ref_out)
    bft(rf, out_as_numpy=False, buffer=outv)
    torch.cuda.synchronize()
    E.lib.pycbf_profile_reset()
    E.lib.pycbf_profile_enable(1)
    t0 = time.perf_counter()
    for _ in range(steps):
        bft(rf, out_as_numpy=False, buffer=outv)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    E.lib.pycbf_profile_enable(0)
    gather_ms, _ = E.profile_read("das_gather")
    err = float((outv - ref_out).abs().max() / ref_out.abs().max())
    del bft, outv
    torch.cuda.empty_cache()
    return ms, err, gather_ms


SLSC_NXO, SLSC_NZO, SLSC_SUB, SLSC_LAGS, SLSC_KWIN = 256, 256, 8, list(range(1, 26)), 8


def slsc_leg(dev, reps, with_cpu):
    """Coherence workload: 75 plane waves x 128 rx beamformed per receive
    channel (cubic, transmits summed) on a 256 x 2048 fine grid over 40 mm of
    depth (speckle with a 5 mm anechoic cyst), then SLSC with lags 1..25 and
    an 8-sample axial kernel onto the 256 x 256 output grid, in FP32 and on
    MFMA with FP16 / BF16 inputs.  Returns the JSON object of the leg."""
    import torch

    from pycbf_amd import sim
    from pycbf_amd.gpu import SyntheticBeamformer, slsc
    from pycbf_amd.gpu import _engine as E
    steers, xele, params, (nx, nzf) = sim.workload_slsc(NRX, NTX, SLSC_NXO, SLSC_NZO, SLSC_SUB)
    xs, zs, amp = sim.cyst_speckle(20_000, (-19e-3, 19e-3), (2e-3, 42e-3), (0.0, 22e-3), 2.5e-3,
                                   4321)
    rf = sim.simulate_rf_torch(steers, xele, NT, xs, zs, amp, dev, seed=4321)
    bf = SyntheticBeamformer(**params, t0=0.0, dt=1.0 / 40e6, nt=NT, interp={"kind": "cubic"},
                             sumtype="tx_only", device=dev)
    imrx = torch.empty((NRX, nx * nzf), dtype=torch.float32, device=dev)
    bf(rf, out_as_numpy=False, buffer=imrx)
    torch.cuda.synchronize()
    das_reps = 3
    t0 = time.perf_counter()
    for _ in range(das_reps):
        bf(rf, out_as_numpy=False, buffer=imrx)
    torch.cuda.synchronize()
    das_ms = (time.perf_counter() - t0) * 1e3 / das_reps
    log(f"slsc leg: per-rx DAS {NTX}x{NRX}x{nx}x{nzf} {das_ms:.3f} ms")
    im4 = imrx.view(1, NRX, nx, nzf)
    from pycbf_amd.gpu import slsc_selection
    sel = slsc_selection(bf.tables()[3].view(NRX, nx, nzf), SLSC_KWIN, -SLSC_KWIN // 2, 1,
                         SLSC_SUB, device=dev)
    out = torch.empty((1, SLSC_NXO, SLSC_NZO), dtype=torch.float32, device=dev)
    names = {"fp32": "slsc_fp32", "fp16": "slsc_mfma_fp16", "bf16": "slsc_mfma_bf16"}
    ms, outs = {}, {}
    for prec, kname in names.items():
        slsc(im4, SLSC_LAGS, SLSC_KWIN, -SLSC_KWIN // 2, None, 1, SLSC_SUB, prec, out=out, sel=sel)
        torch.cuda.synchronize()
        E.lib.pycbf_profile_reset()
        E.lib.pycbf_profile_enable(1)
        for _ in range(reps):
            slsc(im4, SLSC_LAGS, SLSC_KWIN, -SLSC_KWIN // 2, None, 1, SLSC_SUB, prec, out=out, sel=sel)
        torch.cuda.synchronize()
        E.lib.pycbf_profile_enable(0)
        k_ms, k_n = E.profile_read(kname)
        ms[prec] = k_ms / max(k_n, 1)
        outs[prec] = out[0].clone()
        log(f"slsc {prec}: {ms[prec] * 1e3:.1f} us/launch")
    npx = SLSC_NXO * SLSC_NZO
    algo = 4.0 * NRX * nx * nzf + 36.0 * npx     # per-rx images + selection bits + output
    res = {
        "workload": f"SLSC lags 1..25, 8-sample axial kernel, {NRX} rx x {NTX} PW, "
                    f"{SLSC_NXO}x{SLSC_NZO} px from a {nx}x{nzf} per-rx grid (cyst phantom)",
        "px_per_s": {p: round(npx / (ms[p] * 1e-3), 1) for p in ms},
        "kernel_us": {p: round(ms[p] * 1e3, 2) for p in ms},
        "mfma_fp16_speedup_vs_fp32": round(ms["fp32"] / ms["fp16"], 3),
        "mfma_bf16_speedup_vs_fp32": round(ms["fp32"] / ms["bf16"], 3),
        "roofline_fp16": {"bound": "hbm", "achieved": round(algo / (ms["fp16"] * 1e-3) / 1e9, 1),
                          "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                          "frac": round(algo / (ms["fp16"] * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)},
        "per_rx_das_ms": round(das_ms, 3),
        "pipeline_px_per_s": round(npx / ((das_ms + ms["fp16"]) * 1e-3), 1),
    }
    if with_cpu:
        from oracle import numerics
        cols = list(range(0, nx, 16))
        sub = imrx.view(NRX, nx, nzf)[:, cols].cpu().numpy()
        asub = bf.tables()[3].view(NRX, nx, nzf)[:, cols].cpu().numpy()
        t0 = time.perf_counter()
        ref = numerics.slsc_image(sub, asub, SLSC_LAGS, SLSC_KWIN, -SLSC_KWIN // 2, 1, SLSC_SUB)
        wall = time.perf_counter() - t0
        n = len(cols) * SLSC_NZO
        # error of each precision on the sampled columns (coherence in [-1, 1];
        # NaN where the oracle has NaN counts as agreement)
        err = {}
        for prec, o in outs.items():
            g = o[cols].double().cpu().numpy()
            both = np.isnan(g) & np.isnan(ref)
            d = np.where(both, 0.0, np.abs(g - ref))
            err[prec] = float(np.nanmax(np.where(np.isnan(d), np.inf, d)))
        res["max_abs_err_vs_fp64"] = {p: float(f"{e:.3g}") for p, e in err.items()}
        log(f"slsc max abs error vs FP64 on {n} px: {err}")
        res["cpu_baseline"] = {"value": round(n / wall, 1), "unit": "px/s", "cores": 1,
                               "kind": "port",
                               "sample": f"notebook pixel loop (numpy FP64) on {len(cols)} of "
                                         f"{SLSC_NXO} columns ({n} px), {wall:.2f} s"}
        log(f"slsc cpu baseline: {n / wall:.1f} px/s")
    dres = demod_leg(dev, imrx.view(NRX * nx, nzf), 40e-3 / (nzf - 1), reps, with_cpu)
    del rf, imrx, sel
    torch.cuda.empty_cache()
    return res, dres


def demod_leg(dev, lines, dz, reps, with_cpu):
    """Demodulation of the coherence leg's per-receive images to IQ along
    depth (the notebooks' fft_tukey_demod, 5-12 MHz band at 1540 m/s): every
    line of 2048 axial samples of all receive channels and columns, one
    workgroup per line running the real FFT, band, window, shift and the
    Bluestein inverse DFT in LDS (pycbf_demod_fft); the dense operator product
    (three-way BF16 split on the matrix units, pycbf_demod_apply) is timed
    beside it.  Returns the JSON object of the leg."""
    import torch

    from pycbf_amd.gpu import _engine as E
    from pycbf_amd.gpu.demod import band, fft_tukey_demod
    kmin, kmax, ks = 2 * 5e6 / 1540.0, 2 * 12e6 / 1540.0, 1.0 / dz
    rows, n = lines.shape
    ikmin, ikmax = band(n, kmin, kmax, ks)
    m = ikmax - ikmin

    def timed(method, kname):
        iq = fft_tukey_demod(lines, kmin, kmax, ks, method=method)
        torch.cuda.synchronize()
        E.lib.pycbf_profile_reset()
        E.lib.pycbf_profile_enable(1)
        t0 = time.perf_counter()
        for _ in range(reps):
            iq = fft_tukey_demod(lines, kmin, kmax, ks, method=method)
        torch.cuda.synchronize()
        call = (time.perf_counter() - t0) * 1e3 / reps
        E.lib.pycbf_profile_enable(0)
        k_ms, k_n = E.profile_read(kname)
        return iq, k_ms / max(k_n, 1), call

    iq, kms, call_ms = timed("fft", "demod_fft")
    iq_op, op_ms, _ = timed("operator", "demod_gemm")
    # compulsory HBM bytes: every line read once, its IQ written once (the
    # tables, a few hundred KB, stay in cache)
    hbm = 4.0 * rows * n + 8.0 * rows * m
    achieved = hbm / (kms * 1e-3) / 1e9
    op_err = float((iq_op - iq).abs().max() / iq.abs().max())
    log(f"demod: {rows} lines x {n} -> {m} IQ, FFT kernel {kms * 1e3:.1f} us "
        f"({achieved:.0f} GB/s, {achieved / HBM_PEAK_GBPS:.3f} of HBM), operator product "
        f"{op_ms * 1e3:.1f} us (max difference {op_err:.2e} of peak), call {call_ms:.3f} ms")
    res = {
        "workload": f"fft_tukey_demod of {rows} lines x {n} axial samples (dz {dz * 1e6:.2f} um, "
                    f"5-12 MHz) to {m} IQ samples",
        "lines_per_s": round(rows / (kms * 1e-3), 1),
        "kernel": "demod_fft",
        "kernel_us": round(kms * 1e3, 2),
        "call_ms": round(call_ms, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "bytes": "lines read once (float32) + IQ written once (complex64)"},
        "operator_product": {"kernel": "demod_gemm", "kernel_us": round(op_ms * 1e3, 2),
                             "speedup_fft_vs_operator": round(op_ms / kms, 3),
                             "max_diff_rel_peak_vs_fft": float(f"{op_err:.3g}")},
    }
    if with_cpu:
        from oracle import demod as odemod
        sub = lines[:2048].double().cpu().numpy()
        t0 = time.perf_counter()
        ref = odemod.fft_tukey_demod(sub, kmin, kmax, ks)
        wall = time.perf_counter() - t0
        got = iq[:2048].cpu().numpy()
        res["max_abs_err_rel_peak"] = float(f"{np.abs(got - ref).max() / np.abs(ref).max():.3g}")
        res["cpu_baseline"] = {"value": round(sub.shape[0] / wall, 1), "unit": "lines/s",
                               "cores": 1, "kind": "port",
                               "sample": f"numpy FFT chain (FP64) on 2048 of {rows} lines, "
                                         f"{wall:.3f} s"}
        log(f"demod cpu baseline: {sub.shape[0] / wall:.1f} lines/s; error "
            f"{res['max_abs_err_rel_peak']} of peak")
    return res


def sos_leg(dev, steps):
    """Differentiable workload: 31 plane waves over +-15 deg x 128 rx x 4096
    samples, 256 x 256 pixels, a 64 x 64 sound-speed map (1450 m/s above
    10 mm, 1540 m/s below, +-20 m/s blobs); channel data simulated through the
    true map with the straight-ray model.  From a uniform 1540 m/s start the
    64 x 64 map is fitted to the cross-angle consistency of the per-transmit
    images (SoSProblem.angle_loss, lag 3): `steps` / 4 Adam steps on a depth
    profile (one speed per map row, lr 5 m/s), then 3 `steps` / 4 on the free
    map (lr 2 m/s, smoothness 1e-3), each a forward and a backward pass through
    the ray-delay and DAS kernels.  Labelled variants: the image-focus loss on
    the free map and on a 4 x 4 grid of speeds, 50 steps each.  Returns the JSON
    object of the leg."""
    import torch

    from pycbf_amd import sim
    from pycbf_amd.gpu import _engine as E
    from pycbf_amd.gpu.sos import SlownessGrid, SoSProblem, ray_delays
    w = sim.workload_sos_geometry()
    geo = SlownessGrid.covering(w["xlim"], w["zlim"], w["nxm"], w["nzm"])
    f = lambda a: torch.as_tensor(np.asarray(a, np.float32), device=dev)  # noqa: E731
    ctrue = sim.sos_true_map(w["nzm"], w["nxm"], w["xlim"], w["zlim"], seed=11)
    xs, zs, amp = sim.scatterers(50, 10_000, w["xlim"], (w["zlim"][0], w["zlim"][1]), 11)
    ttx, trx = ray_delays(1.0 / f(ctrue), geo, f(w["steers"]), f(w["xorig"]), f(w["xele"]),
                          f(np.stack([xs, zs], 1)), 1540.0, 64)
    rf = sim.simulate_rf_delays_torch(ttx, trx, amp, w["nt"], dev, seed=11)
    prob = SoSProblem(rf=rf, steer=f(w["steers"]), xorig=f(w["xorig"]), xele=f(w["xele"]),
                      pvec=f(w["pnts"]), gridshape=w["gridshape"], geo=geo, t0=0.0,
                      dt=1.0 / sim.FS)
    ct = f(ctrue)
    c0 = torch.full_like(ct, 1540.0)
    top = int(round((10e-3 - geo.zm0) / geo.dzm))      # map rows above the 10 mm layer boundary
    with torch.no_grad():
        e0 = float((prob.images(c0).sum(0) ** 2).mean())
        angle_true = float(prob.angle_loss(ct))
        angle_start = float(prob.angle_loss(c0))
        focus_true = float(prob.loss(ct, e0))

    def optimise(theta, cmap, lossfn, lr, smooth, nsteps, timed=False):
        prob.smooth_weight = smooth
        opt = torch.optim.Adam([theta], lr=lr)
        hist = []
        if timed:
            E.lib.pycbf_profile_reset()
            E.lib.pycbf_profile_enable(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            opt.zero_grad()
            c = cmap()
            loss = lossfn(c)
            loss.backward()
            opt.step()
            # kept on the device: reading here would stall the host every step
            hist.append(torch.stack([loss.detach().float(),
                                     torch.sqrt(((c.detach() - ct) ** 2).mean())]))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        if timed:
            E.lib.pycbf_profile_enable(0)
        prob.smooth_weight = 0.0
        return wall, torch.stack(hist).cpu().numpy(), cmap().detach()

    angle = lambda c: prob.angle_loss(c, 3)  # noqa: E731
    nprof = max(1, steps // 4)
    nfree = max(1, steps - nprof)
    prof = torch.full((w["nzm"], 1), 1540.0, device=dev, requires_grad=True)
    _, hist_p, cprof = optimise(prof, lambda: prof.expand(w["nzm"], w["nxm"]), angle, 5.0, 0.0,
                                nprof)
    free = cprof.clone().requires_grad_()
    wall, hist_f, cfin = optimise(free, lambda: free, angle, 2.0, 1e-3, nfree, timed=True)
    kern = {}
    for k in ("ray_delays", "ray_delays_bwd", "das_gather", "das_bwd_tau", "spline_coefs"):
        ms, n = E.profile_read(k)
        kern[k] = round(ms / max(n, 1), 4)
    with torch.no_grad():
        angle_fin = float(prob.angle_loss(cfin))
    rmse = lambda c: float(torch.sqrt(((c - ct) ** 2).mean()))  # noqa: E731
    layers = lambda c: (round(float(c[:top].mean()), 2), round(float(c[top:].mean()), 2))  # noqa: E731

    # labelled variants: the image-focus loss (free map; 4 x 4 grid of speeds)
    focus = lambda c: prob.loss(c, e0)  # noqa: E731
    ffree = c0.clone().requires_grad_()
    _, hist_ff, cff = optimise(ffree, lambda: ffree, focus, 2.0, 1e-4, 50)
    theta = torch.full((1, 1, 4, 4), 1540.0, device=dev, requires_grad=True)

    def coarse():
        return torch.nn.functional.interpolate(theta, size=(w["nzm"], w["nxm"]), mode="bilinear",
                                               align_corners=True)[0, 0]

    _, hist4, c4 = optimise(theta, coarse, focus, 10.0, 0.0, 50)
    ntx, nrx = len(w["steers"]), len(w["xele"])
    samples = ntx * nrx * w["gridshape"][0] * w["gridshape"][1]
    log(f"sos leg: cross-angle fit, {nprof} depth-profile + {nfree} free-map Adam steps "
        f"({wall * 1e3 / nfree:.2f} ms/step): loss {angle_start:.4f} -> {angle_fin:.4f} (true map "
        f"{angle_true:.4f}), map rmse {rmse(c0):.2f} -> {rmse(cprof):.2f} -> {rmse(cfin):.2f} m/s, "
        f"layers {layers(c0)} -> {layers(cfin)} (true {layers(ct)}) m/s; focus loss, free map: "
        f"rmse -> {rmse(cff):.2f}; 4x4: rmse -> {rmse(c4):.2f}")
    del rf, prob
    torch.cuda.empty_cache()
    r3 = lambda a: [round(float(v), 3) for v in a]  # noqa: E731
    r5 = lambda a: [round(float(v), 5) for v in a]  # noqa: E731
    return {
        "workload": f"differentiable DAS for speed of sound: {ntx} PW x {nrx} rx x {w['nt']} "
                    f"samples, {w['gridshape'][0]}x{w['gridshape'][1]} px, "
                    f"{w['nzm']}x{w['nxm']} map from 1540 m/s fitted to the cross-angle "
                    f"consistency of the per-transmit images (lag 3): {nprof} Adam steps on a "
                    f"depth profile (lr 5 m/s), {nfree} on the free map (lr 2 m/s, smoothness "
                    f"1e-3); straight-ray delays with float32 residuals, cubic",
        "ms_per_step": round(wall * 1e3 / nfree, 3),
        "fwd_bwd_gsamples_per_s": round(samples / (wall / nfree) / 1e9, 2),
        "loss_per_step": r5(np.concatenate([hist_p[:, 0], hist_f[:, 0]])),
        "map_rmse_per_step_m_per_s": r3(np.concatenate([hist_p[:, 1], hist_f[:, 1]])),
        "loss_start": round(angle_start, 6), "loss_final": round(angle_fin, 6),
        "loss_at_true_map": round(angle_true, 6),
        "map_rmse_m_per_s": {"start": round(rmse(c0), 3), "after_profile": round(rmse(cprof), 3),
                             "final": round(rmse(cfin), 3)},
        "layer_means_m_per_s": {"start": layers(c0), "final": layers(cfin), "true": layers(ct),
                                "rows": "above / below 10 mm"},
        "kernel_ms": kern,
        "variant_focus_loss_free_map": {
            "lr": 2.0, "smoothness": 1e-4, "steps": 50, "loss_at_true_map": round(focus_true, 6),
            "loss_final": round(float(hist_ff[-1, 0]), 6),
            "map_rmse_m_per_s": [round(rmse(c0), 3), round(rmse(cff), 3)],
            "loss_per_step": r5(hist_ff[:, 0])},
        "variant_focus_loss_4x4_grid": {
            "lr": 10.0, "steps": 50, "map_rmse_m_per_s": [round(rmse(c0), 3), round(rmse(c4), 3)],
            "loss_per_step": r5(hist4[::5, 0])},
    }


def adjoint_leg(dev, reps):
    """Channel-data adjoint: the cubic Tabbed DAS of 31 plane waves over
    +-15 deg x 128 rx x 4096 samples onto 256 x 256 pixels (the
    differentiable workload's shape), forward and backward with respect to the
    channel data through the autograd Function; seeded Gaussian channel data
    and upstream gradient.  Returns the JSON object of the leg."""
    import torch

    from pycbf_amd import sim
    from pycbf_amd.gpu import _engine as E
    from pycbf_amd.gpu.autograd import das_tabbed
    ntx, nrx, nt, side = 31, 128, 4096, 256
    xele = sim.linear_array(nrx)
    steers = np.radians(np.linspace(-15, 15, ntx))
    _, _, pnts = sim.grid(side, side, (-19e-3, 19e-3), (5e-3, 45e-3))
    tabs = [torch.as_tensor(a, device=dev) for a in sim.pw_tabs(steers, xele, pnts)]
    g = torch.Generator(device=dev)
    g.manual_seed(18)
    rf = torch.randn((ntx, nrx, nt), generator=g, device=dev, requires_grad=True)
    G = torch.randn((side * side,), generator=g, device=dev)

    def step():
        out = das_tabbed(rf, *tabs, 0.0, 1.0 / sim.FS, "cubic", "tx_and_rx", 0.0)
        (grf,) = torch.autograd.grad((out * G).sum(), [rf])
        return grf

    step()
    torch.cuda.synchronize()
    E.lib.pycbf_profile_reset()
    E.lib.pycbf_profile_enable(1)
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    E.lib.pycbf_profile_enable(0)
    kern = {}
    for k in ("das_gather", "das_bwd_rf", "spline_coefs"):
        ms, n = E.profile_read(k)
        kern[k] = round(ms / max(n, 1), 4)
    samples = ntx * nrx * side * side
    log(f"adjoint leg: forward + channel-data backward {wall * 1e3:.3f} ms, "
        f"das_bwd_rf {kern['das_bwd_rf']:.3f} ms")
    del rf, tabs
    torch.cuda.empty_cache()
    syn = synthetic_adjoint_leg(dev, reps, ntx, nrx, nt, side)
    return {
        "workload": f"cubic Tabbed DAS {ntx} PW x {nrx} rx x {nt} samples onto {side}x{side} px, "
                    "forward + backward with respect to the channel data",
        "ms_fwd_bwd": round(wall * 1e3, 3),
        "fwd_bwd_gsamples_per_s": round(samples / wall / 1e9, 2),
        "kernel_ms": kern,
        "synthetic": syn,
    }


def synthetic_adjoint_leg(dev, reps, ntx, nrx, nt, side):
    """The same shape as a Synthetic DAS (plane waves as synthetic points):
    forward + backward with respect to the channel data and c0 through
    das_synthetic (geometry re-evaluated per tile in registers, no table),
    against das_tabbed on the beamformer's stored tables (evaluated once,
    outside the timed step).  Returns the JSON object."""
    import torch

    from pycbf_amd import sim
    from pycbf_amd.gpu import SyntheticBeamformer
    from pycbf_amd.gpu.autograd import das_synthetic, das_tabbed
    xele = sim.linear_array(nrx)
    steers = np.radians(np.linspace(-15, 15, ntx))
    _, _, pnts = sim.grid(side, side, (-19e-3, 19e-3), (5e-3, 45e-3))
    bf = SyntheticBeamformer(**sim.pw_synthetic(steers, xele, pnts), t0=0.0, dt=1.0 / sim.FS,
                             nt=nt, interp={"kind": "cubic"}, sumtype="tx_and_rx", device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(28)
    rf = torch.randn((1, ntx, nrx, nt), generator=g, device=dev, requires_grad=True)
    G = torch.randn((1, side * side), generator=g, device=dev)
    c0 = torch.tensor(float(bf.c0), dtype=torch.float32, requires_grad=True)
    tabs = bf.tables()
    table_bytes = sum(t.numel() * 4 for t in tabs)

    def step_syn():
        out = das_synthetic(rf, bf)
        return torch.autograd.grad((out * G).sum(), [rf])

    def step_syn_c0():
        out = das_synthetic(rf, bf, c0)
        return torch.autograd.grad((out * G).sum(), [rf, c0])

    def step_tab():
        out = das_tabbed(rf, *tabs, 0.0, 1.0 / sim.FS, "cubic", "tx_and_rx", float(bf.thresh))
        return torch.autograd.grad((out * G).sum(), [rf])

    res = {}
    for name, fn in (("registers", step_syn), ("registers_c0", step_syn_c0),
                     ("stored_tables", step_tab)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / reps * 1e3
    log(f"synthetic adjoint, fwd+bwd: channel data with geometry in registers "
        f"{res['registers']:.3f} ms (with c0 {res['registers_c0']:.3f} ms); with stored tables "
        f"{res['stored_tables']:.3f} ms; tables {table_bytes / 1e6:.1f} MB")
    del rf, tabs, bf
    torch.cuda.empty_cache()
    return {
        "workload": f"cubic Synthetic DAS {ntx} PW (synthetic points) x {nrx} rx x {nt} samples "
                    f"onto {side}x{side} px, forward + backward",
        "ms_fwd_bwd_registers": round(res["registers"], 3),
        "ms_fwd_bwd_registers_with_c0": round(res["registers_c0"], 3),
        "ms_fwd_bwd_stored_tables": round(res["stored_tables"], 3),
        "table_bytes_avoided": int(table_bytes),
        "grads": "channel data (registers, stored_tables); channel data and c0 "
                 "(registers_with_c0)",
    }


def batch_leg(dev, runs, speckle, measure_peaks=True):
    """Ragged batch: 64 seeded requests, grid side in {128, 192, 256, 384, 512},
    Ntx in {1, 11, 31, 75}, Nrx in {64, 128, 192, 256}, Nt = 4096; cubic DAS
    for 48 requests and SLSC (lags 1..20) for 16; one persistent gather launch
    per run, autotuned tile / receive chunk, the run captured in a HIP graph
    and replayed.  Returns the JSON object of the leg."""
    import torch

    from pycbf_amd import sim
    from pycbf_amd.gpu import _engine as E
    from pycbf_amd.gpu.batch import BatchRequest, RaggedBatch
    rng = np.random.default_rng(2024)
    reqs, specs = [], []
    kinds = ["das"] * 48 + ["slsc"] * 16
    rng.shuffle(kinds)
    for k in range(64):
        side = int(rng.choice([128, 192, 256, 384, 512]))
        ntx = int(rng.choice([1, 11, 31, 75]))
        nrx = int(rng.choice([64, 128, 192, 256]))
        xele = sim.linear_array(nrx)
        steers = np.radians(np.linspace(-16, 16, ntx)) if ntx > 1 else np.zeros(1)
        _, _, pnts = sim.grid(side, side, (-19e-3, 19e-3), (2e-3, 50e-3))
        reqs.append(BatchRequest(params=sim.pw_synthetic(steers, xele, pnts),
                                 gridshape=(side, side), kind=kinds[k], lags=tuple(range(1, 21)),
                                 kwin=5, precision="fp16"))
        specs.append((steers, xele, ntx, nrx, side))
    batch = RaggedBatch(reqs, nt=NT, dt=1.0 / 40e6, device=dev)
    for k, (steers, xele, ntx, nrx, side) in enumerate(specs):
        xs, zs, amp = sim.scatterers(50, speckle, (-19e-3, 19e-3), (2e-3, 50e-3), 5000 + k)
        batch.rf_views[k].copy_(sim.simulate_rf_torch(steers, xele, NT, xs, zs, amp, dev,
                                                      seed=5000 + k))
    torch.cuda.synchronize()
    log(f"batch leg: 64 requests, {batch.samples / 1e9:.2f} G samples per run")
    tile, rchunk, grid = batch.autotune(reps=2)
    log(f"autotune: tile {tile} rchunk {rchunk} ({grid[(tile, rchunk)]:.2f} ms gather)")
    batch.capture()
    batch.run()
    torch.cuda.synchronize()
    E.lib.pycbf_profile_reset()
    t0 = time.perf_counter()
    for _ in range(runs):
        batch.run()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / runs
    lat = batch.latencies_ms()
    log(f"batch: {ms:.2f} ms/run, {batch.samples / (ms * 1e-3) / 1e9:.1f} Gsamples/s, latency p50 "
        f"{np.percentile(lat, 50):.2f} p99 {np.percentile(lat, 99):.2f} ms")
    # HBM bound of the gather on its compulsory bytes, per request: the
    # (y, h*y2) table once (8 B per entry, nt + 1 entries per channel), the
    # pixel positions, and the output written once (summed image, or the
    # per-receive images of an SLSC request).  The 16 B stencil of every
    # contributing sample comes from L1/L2 (the batch kernel gathers directly,
    # no LDS staging): the on-chip bound is that traffic against the measured
    # L2 read bandwidth.
    from pycbf_amd.gpu import SyntheticBeamformer
    hbm = 0.0
    contrib = 0
    for (steers, xele, ntx, nrx, side), req, kind in zip(specs, reqs, kinds):
        npx = side * side
        hbm += (8.0 * ntx * nrx * (NT + 1) + 8.0 * npx
                + 4.0 * npx * (nrx if kind == "slsc" else 1))
        bfr = SyntheticBeamformer(**req.params, t0=0.0, dt=1.0 / 40e6, nt=NT,
                                  interp={"kind": "cubic"}, sumtype="tx_and_rx", device=dev)
        contrib += count_contributing(bfr, req.thresh)
        del bfr
    g_ms = grid[(tile, rchunk)]
    g_achieved = hbm / (g_ms * 1e-3) / 1e9
    l2_bytes = 16.0 * contrib
    l2_achieved = l2_bytes / (g_ms * 1e-3) / 1e9
    l2_peak = E.measure_bandwidth("l2", dev) if measure_peaks else None
    log(f"batch gather: {hbm / 1e9:.2f} GB compulsory, {g_achieved:.0f} GB/s "
        f"({g_achieved / HBM_PEAK_GBPS:.3f} of HBM); stencil {l2_bytes / 1e9:.1f} GB from cache "
        f"({contrib / batch.samples:.3f} of samples contribute), {l2_achieved:.0f} GB/s"
        + (f" ({l2_achieved / l2_peak:.3f} of measured L2)" if l2_peak else ""))
    res = {
        "workload": "ragged batch of 64 seeded requests (48 cubic DAS, 16 SLSC lags 1..20), grids "
                    "128..512 square, Ntx 1..75, Nrx 64..256, Nt 4096; one persistent gather "
                    "launch per run, HIP-graph replay",
        "ms_per_run": round(ms, 3),
        "gsamples_per_s": round(batch.samples / (ms * 1e-3) / 1e9, 2),
        "samples_per_run": int(batch.samples),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3),
                       "max": round(float(lat.max()), 3)},
        "roofline": {"bound": "hbm", "achieved": round(g_achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(g_achieved / HBM_PEAK_GBPS, 4),
                     "kernel": "batch_gather", "kernel_ms": round(g_ms, 3),
                     "bytes": "spline tables read once + pixel positions + outputs"},
        "roofline_onchip": {"bound": "l2", "achieved": round(l2_achieved, 1),
                            "peak": round(l2_peak, 1) if l2_peak else None, "unit": "GB/s",
                            "frac": round(l2_achieved / l2_peak, 4) if l2_peak else None,
                            "bytes": "16 B per contributing sample",
                            "contributing_samples": int(contrib)},
        "autotune": {"tile": tile, "rchunk": rchunk,
                     "gather_ms": {f"{t}x{t}/{c}": round(v, 3) for (t, c), v in grid.items()}},
        "data": f"seeded channel data, 50 points + {speckle} speckle scatterers per request",
    }
    del batch
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=1)
    ap.add_argument("--cpu-frames", type=int, default=12,
                    help="frames timed for the CPU baseline (about 10-30 s of host work)")
    ap.add_argument("--no-slsc", action="store_true", help="skip the coherence leg")
    ap.add_argument("--no-sos", action="store_true", help="skip the differentiable leg")
    ap.add_argument("--sos-steps", type=int, default=400)
    ap.add_argument("--no-batch", action="store_true", help="skip the ragged-batch leg")
    ap.add_argument("--batch-runs", type=int, default=5)
    ap.add_argument("--batch-speckle", type=int, default=10_000)
    ap.add_argument("--slsc-reps", type=int, default=20)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the channel-data adjoint leg")
    ap.add_argument("--adjoint-reps", type=int, default=10)
    ap.add_argument("--data", choices=["sim", "noise"], default="sim",
                    help="noise: seeded Gaussian RF of the same shape (profiling runs)")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the plain-GPU comparators and the bandwidth measurements "
                         "(profiling runs)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one elementwise copy of this many MiB before timing "
                         "(reference kernel for scaling FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}, WORLD_SIZE {world}); "
              "the hot path shards over frames / pixels with no exchange and is not run "
              "across GPUs here", file=sys.stderr)
        sys.exit(2)

    import torch

    from pycbf_amd.gpu import SyntheticBeamformer
    from pycbf_amd.gpu import _engine as E

    dev = torch.device("cuda", 0)
    xele, steers, pnts, params = workload_geometry()
    log(f"workload: {NTX} tx x {NRX} rx x {NX}x{NZ} px, nt={NT}, {args.frames} frames")
    if args.data == "sim":
        rf = make_rf(steers, xele, args.frames, dev)
    else:
        g = torch.Generator(device=dev)
        g.manual_seed(1234)
        rf = torch.randn((args.frames, NTX, NRX, NT), generator=g, device=dev)
    if args.calib_copy_mib > 0:
        n = args.calib_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        # evict the freshly written source from L2 and the infinity cache, so
        # that the copy's reads all come from HBM
        flush = torch.ones(2 * n, dtype=torch.float32, device=dev)
        del flush
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
        log(f"calibration copy: {args.calib_copy_mib} MiB read + written")
    bf = SyntheticBeamformer(**params, t0=0.0, dt=1.0 / 40e6, nt=NT, interp={"kind": "cubic"},
                             sumtype="tx_and_rx", device=dev)
    out = torch.empty((args.frames, NX * NZ), dtype=torch.float32, device=dev)
    for i in range(args.warmup):
        bf(rf, out_as_numpy=False, buffer=out)
        torch.cuda.synchronize()
        log(f"warmup {i + 1}/{args.warmup}")
    E.lib.pycbf_profile_reset()
    E.lib.pycbf_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bf(rf, out_as_numpy=False, buffer=out)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    E.lib.pycbf_profile_enable(0)
    ms_step = (t1 - t0) * 1e3 / args.steps
    samples = NTX * NRX * NX * NZ * args.frames
    value = samples / (ms_step * 1e-3) / 1e9
    log(f"{ms_step:.3f} ms/step, {value:.1f} Gsamples/s")

    g_ms, g_n = E.profile_read("das_gather")
    s_ms, s_n = E.profile_read("spline_coefs")
    t_ms, t_n = E.profile_read("synthetic_tables")
    gather_ms = g_ms / max(g_n, 1)
    npx = NX * NZ
    log(f"gather {gather_ms:.3f} ms/launch, spline {s_ms / max(s_n, 1):.3f} ms, "
        f"tables {t_ms / max(t_n, 1):.3f} ms")
    # HBM roofline of the gather on its compulsory bytes: the (y, h*y2) spline
    # table read once (8 B per entry, nt + 1 entries per channel), the pixel
    # positions and processing order, the images written
    table_bytes = 8.0 * args.frames * NTX * NRX * (NT + 1)
    compulsory = table_bytes + 12.0 * npx + 4.0 * npx * args.frames
    hbm_achieved = compulsory / (gather_ms * 1e-3) / 1e9
    prof = profile_figures(args.frames)
    traffic = prof.get("traffic")
    # on-chip: the stencil reads (16 B per contributing sample and frame) the
    # gather serves from LDS, against the measured LDS read bandwidth
    contrib = count_contributing(bf, 1e-2)
    lds_bytes = 16.0 * contrib * args.frames
    lds_achieved = lds_bytes / (gather_ms * 1e-3) / 1e9
    peaks = {"hbm": None, "l2": None, "lds": None}
    plain_ms = plain_err = direct_ms = direct_err = tab_ms = tab_err = tab_gather = None
    if not args.no_compare:
        peaks = {k: round(E.measure_bandwidth(k, dev), 1) for k in ("hbm", "l2", "lds")}
        log(f"measured read bandwidth GB/s: {peaks}; contributing samples per frame {contrib} "
            f"({contrib / (NTX * NRX * npx):.3f} of all)")
        # plain-GPU comparators on the same workload and data
        plain_ms, plain_err = time_variant("PYCBF_DAS_PLAIN", "1", params, rf, out, 2)
        direct_ms, direct_err = time_variant("PYCBF_DAS_LDS", "0", params, rf, out, 2)
        tab_ms, tab_err, tab_gather = time_tabbed(bf, rf, out, 3)
        tab_gather /= 3
        log(f"TabbedBeamformer on the same tables {tab_ms:.2f} ms/step, gather "
            f"{tab_gather:.3f} ms (err {tab_err:.2e})")
        log(f"thread-per-sample comparator {plain_ms:.2f} ms/step (err {plain_err:.2e}); direct "
            f"gather {direct_ms:.2f} ms/step (err {direct_err:.2e})")

    cpu = parity = None
    if not args.no_cpu:
        info = cpu_info()
        threads = min(16, info["usable"])
        nf = min(args.cpu_frames, args.frames)
        frames = [rf[f].cpu().numpy() for f in range(nf)]
        rate, used, kind, desc, images = cpu_baseline(frames, params, args.cpu_stride, threads)
        rate1, _, _, desc1, _ = cpu_baseline(frames[:1], params, 16, 1)
        cpu = {"value": round(rate, 4), "unit": "Gsamples/s", "cores": used, "kind": kind,
               "sample": desc, "cpu_model": info["model"], "cpu_count": info["cpu_count"],
               "cpus_usable": info["usable"],
               "one_core": {"value": round(rate1, 4), "unit": "Gsamples/s", "sample": desc1}}
        log(f"cpu baseline: {rate:.4f} Gsamples/s on {used} threads ({kind}); one core "
            f"{rate1:.4f}")
        # parity of the timed GPU frames with the CPU images of the same frames
        sub = torch.arange(0, npx, args.cpu_stride, device=dev)
        gpu = out[:nf, sub].double().cpu().numpy()
        errs = [float(np.abs(gpu[f] - images[f]).max() / np.abs(images[f]).max())
                for f in range(nf)]
        parity = {"frames": nf, "pixels_per_frame": int(sub.numel()),
                  "max_err_rel_peak": float(f"{max(errs):.3g}"), "bar": 1e-5,
                  "against": f"cpu_baseline images ({kind} C)"}
        log(f"parity: max error {max(errs):.3g} of the peak over {nf} frames")
        del frames

    slsc_res = demod_res = None
    if not args.no_slsc:
        del rf
        torch.cuda.empty_cache()
        slsc_res, demod_res = slsc_leg(dev, args.slsc_reps, not args.no_cpu)
    adj_res = None if args.no_adjoint else adjoint_leg(dev, args.adjoint_reps)
    sos_res = None if args.no_sos else sos_leg(dev, args.sos_steps)
    batch_res = None if args.no_batch else batch_leg(dev, args.batch_runs, args.batch_speckle,
                                                     not args.no_compare)

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded plane-wave channel data: 50 points + 1e4 speckle per frame)",
        "config": {"workload": f"Synthetic cubic DAS {NTX}tx x {NRX}rx x {NX}x{NZ}px, "
                               f"nt={NT}, {args.frames} frames, sumtype tx_and_rx",
                   "ntx": NTX, "nrx": NRX, "nt": NT, "grid": [NX, NZ], "frames": args.frames,
                   "interp": "cubic"},
        "roofline": {"bound": "hbm", "achieved": round(hbm_achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(hbm_achieved / HBM_PEAK_GBPS, 4),
                     "traffic": traffic, "kernel": "das_gather",
                     "kernel_ms": round(gather_ms, 4), "compulsory_bytes": compulsory,
                     "bytes": "spline table read once + pixel positions and order + images"},
        "roofline_onchip": {"bound": "lds", "achieved": round(lds_achieved, 1),
                            "peak": peaks["lds"], "unit": "GB/s",
                            "frac": round(lds_achieved / peaks["lds"], 4) if peaks["lds"] else None,
                            "bytes": "16 B per contributing sample and frame",
                            "contributing_samples_per_frame": contrib,
                            "all_samples_per_frame": NTX * NRX * npx},
        "roofline_issue": None if not prof.get("valu") else {
            "bound": "valu", "unit": "Ginstr/s",
            "achieved": round(prof["valu"] / (gather_ms * 1e-3) / 1e9, 1),
            "peak": VALU_PEAK_GINSTR,
            "frac": round(prof["valu"] / (gather_ms * 1e-3) / 1e9 / VALU_PEAK_GINSTR, 4),
            "instructions": prof["valu"],
            "bytes": "wave-level VALU instructions per launch (SQ_INSTS_VALU, committed "
                     "profile) over this run's gather time; peak one float32 wave64 VALU "
                     "instruction per CU per cycle at the 2.4 GHz peak clock"},
        "peaks_measured_gbps": peaks,
        "plain_gpu": None if plain_ms is None else {
            "value": round(samples / (plain_ms * 1e-3) / 1e9, 2), "unit": "Gsamples/s",
            "ms_per_step": round(plain_ms, 3),
            "max_err_rel_peak_vs_headline": float(f"{plain_err:.3g}"),
            "kernel": "thread per (frame, tx, rx, px) sample with one float atomic each: "
                      "pycbf's GPU launch scheme (PYCBF_DAS_PLAIN=1)"},
        "tabbed_gpu": None if tab_ms is None else {
            "value": round(samples / (tab_ms * 1e-3) / 1e9, 2), "unit": "Gsamples/s",
            "ms_per_step": round(tab_ms, 3), "gather_ms": round(tab_gather, 3),
            "max_err_rel_peak_vs_headline": float(f"{tab_err:.3g}"),
            "kernel": "TabbedBeamformer over the geometry's delay / apodisation tables "
                      "(table-derived pixel order)"},
        "direct_gpu": None if direct_ms is None else {
            "value": round(samples / (direct_ms * 1e-3) / 1e9, 2), "unit": "Gsamples/s",
            "ms_per_step": round(direct_ms, 3),
            "max_err_rel_peak_vs_headline": float(f"{direct_err:.3g}"),
            "kernel": "lane per pixel over every channel, delay tables, direct global "
                      "gathers (PYCBF_DAS_LDS=0)"},
        "speedup_vs_plain_gpu": None if plain_ms is None else round(plain_ms / ms_step, 2),
        "cpu_baseline": cpu,
        "parity": parity,
        "phases_ms": {"das_gather": round(gather_ms, 4),
                      "spline_coefs": round(s_ms / max(s_n, 1), 4),
                      "synthetic_tables": round(t_ms / max(t_n, 1), 4)},
        "slsc": slsc_res,
        "demod": demod_res,
        "adjoint": adj_res,
        "sos": sos_res,
        "batch": batch_res,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
