"""GPU parity at the sizes of the coherence, ragged-batch and speed-of-sound
workloads, against the oracle.

* Coherence workload: 75 plane waves over +-16 deg x 128 receive elements x
  4096 samples beamformed per receive channel (cubic, transmits summed) onto
  a 256 x 2048 grid over 40 mm (speckle with a 5 mm anechoic cyst); SLSC with
  lags 1..25 and an 8-sample axial window onto 256 x 256, checked on a 64 x 64
  subgrid against the FP64 notebook restatement (FP32 within 1e-4, MFMA FP16 /
  BF16 within 1e-3); DMAS over all 128 receive channels against the FP64
  restatement.
* Ragged batch: eight requests at nt = 4096 that cover every grid side
  (128..512), transmit count (1..75) and receive count (64..256) of the
  workload, cubic DAS and SLSC (lags 1..20) in one persistent launch, each
  request against the oracle, for three (tile, receive chunk) settings.
* Speed of sound: 31 plane waves over +-15 deg x 128 rx x 4096 samples onto
  256 x 256 pixels with a 64 x 64 sound-speed map; the end-to-end map gradient
  (ray adjoint + DAS adjoint, float32) against FP64 central differences
  (h = 1e-3 m/s) of the same pipeline at 32 seeded map entries.
"""
import numpy as np
import pytest

import oracle
from oracle import numerics
from pycbf_amd import sim
from tests.helpers import assert_close_to_peak, float32_inputs
from tests.test_gpu_slsc import TOL, _compare

pytestmark = pytest.mark.gpu

NT = 4096
DT = 1.0 / sim.FS
SLSC_LAGS = list(range(1, 26))


# ------------------------------------------------------------- coherence ---
@pytest.fixture(scope="module")
def coh():
    import torch
    from pycbf_amd.gpu import SyntheticBeamformer
    dev = torch.device("cuda", 0)
    steers, xele, params, (nx, nzf) = sim.workload_slsc(128, 75, 256, 256, 8)
    xs, zs, amp = sim.cyst_speckle(20_000, (-19e-3, 19e-3), (2e-3, 42e-3), (0.0, 22e-3), 2.5e-3,
                                   4321)
    rf = sim.simulate_rf_torch(steers, xele, NT, xs, zs, amp, dev, seed=4321)
    bf = SyntheticBeamformer(**params, t0=0.0, dt=DT, nt=NT, interp={"kind": "cubic"},
                             sumtype="tx_only", device=dev)
    imrx = bf(rf, out_as_numpy=False)
    apodrx = bf.tables()[3]
    torch.cuda.synchronize()
    return dict(rf=rf, params=params, nx=nx, nzf=nzf, imrx=imrx, apodrx=apodrx, dev=dev)


def test_coherence_per_receive_images_match_oracle(coh):
    """The per-receive images (one frame, transmits summed) on four columns."""
    print("coherence workload: per-receive images 75x128x4096 onto 256x2048", flush=True)
    nx, nzf = coh["nx"], coh["nzf"]
    cols = np.array([0, 85, 170, 255])
    pix = (cols[:, None] * nzf + np.arange(nzf)[None, :]).ravel()
    p = coh["params"]
    tabs = oracle.synthetic_tables(p["ovectx"], p["nvectx"], p["t0tx"], p["alatx"], p["doftx"],
                                   p["ovecrx"], p["nvecrx"], p["alarx"], p["c0"], p["pnts"][pix])
    ref = oracle.das_tabbed(coh["rf"].cpu().numpy(), 0.0, DT, *tabs, kind="cubic",
                            sumtype="tx_only")
    import torch
    got = coh["imrx"][:, torch.as_tensor(pix, device=coh["dev"])].cpu().numpy()
    assert_close_to_peak(got, ref, 1e-5, "per-receive images")


@pytest.fixture(scope="module")
def slsc_ref(coh):
    nx, nzf = coh["nx"], coh["nzf"]
    cols = np.arange(0, nx, 4)
    im = coh["imrx"].view(128, nx, nzf)[:, cols].cpu().numpy()
    ap = coh["apodrx"].view(128, nx, nzf)[:, cols].cpu().numpy()
    # output rows every 4th of 256: fine pixel 32 k, window 32 k - 4 .. 32 k + 3
    ref = numerics.slsc_image(im, ap, SLSC_LAGS, 8, -4, 1, 32)
    assert ref.shape == (64, 64)
    return cols, ref


@pytest.mark.parametrize("precision", ["fp32", "fp16", "bf16"])
def test_coherence_slsc_subgrid_matches_fp64(coh, slsc_ref, precision):
    import torch
    from pycbf_amd.gpu import slsc, slsc_selection
    print(f"coherence workload: SLSC lags 1..25, K=8, 256x256, {precision}", flush=True)
    nx, nzf = coh["nx"], coh["nzf"]
    sel = slsc_selection(coh["apodrx"].view(128, nx, nzf), 8, -4, 1, 8, device=coh["dev"])
    out = slsc(coh["imrx"].view(1, 128, nx, nzf), SLSC_LAGS, 8, -4, None, 1, 8, precision,
               sel=sel)
    torch.cuda.synchronize()
    assert tuple(out.shape) == (1, 256, 256)
    cols, ref = slsc_ref
    got = out[0].cpu().numpy()[cols][:, ::4]
    _compare(got, ref, TOL[precision], f"coherence 64x64 subgrid {precision}")
    assert np.count_nonzero(np.nan_to_num(ref)) > 0.5 * ref.size


@pytest.mark.parametrize("dmastype", ["ssr", "power", "pssr"])
def test_coherence_dmas_all_receive_channels(coh, dmastype):
    import torch
    from pycbf_amd.gpu import _engine as E
    print(f"coherence workload: DMAS {dmastype} over 128 rx, lags 1..127", flush=True)
    nx, nzf = coh["nx"], coh["nzf"]
    npx = nx * nzf
    lags = np.arange(1, 128, dtype=np.int32)
    dl = torch.as_tensor(lags, device=coh["dev"])
    out = torch.empty(npx, dtype=torch.float32, device=coh["dev"])
    rc = E.lib.pycbf_dmas(E.ptr(coh["imrx"]), 1, 128, npx, E.ptr(dl), lags.size,
                          E.DMASTYPE[dmastype], 1, E.ptr(out), E.stream_of(coh["dev"]))
    E.check(rc, "pycbf_dmas")
    sub = np.arange(0, npx, 61)
    got = out.cpu().numpy()[sub]
    ref = numerics.dmas(coh["imrx"].cpu().numpy()[:, sub], lags, dmastype, True)
    assert_close_to_peak(got, ref, 1e-4, f"DMAS {dmastype}")


# ---------------------------------------------------------- ragged batch ---
BATCH = [  # grid side, ntx, nrx, kind, SLSC precision
    (128, 1, 64, "das", None), (192, 11, 128, "das", None), (256, 31, 192, "slsc", "fp16"),
    (384, 75, 256, "das", None), (512, 11, 64, "das", None), (128, 75, 192, "slsc", "fp32"),
    (256, 1, 256, "das", None), (384, 31, 128, "slsc", "bf16"),
]


@pytest.fixture(scope="module")
def ragged():
    import torch
    from pycbf_amd.gpu.batch import BatchRequest, RaggedBatch
    dev = torch.device("cuda", 0)
    reqs, geos = [], []
    for k, (side, ntx, nrx, kind, prec) in enumerate(BATCH):
        xele = sim.linear_array(nrx)
        steers = np.radians(np.linspace(-16, 16, ntx)) if ntx > 1 else np.zeros(1)
        _, _, pnts = sim.grid(side, side, (-19e-3, 19e-3), (2e-3, 50e-3))
        params = sim.pw_synthetic(steers, xele, pnts)
        reqs.append(BatchRequest(params=params, gridshape=(side, side), kind=kind,
                                 lags=tuple(range(1, 21)), kwin=5,
                                 precision=prec or "fp16"))
        geos.append((steers, xele, params))
    batch = RaggedBatch(reqs, nt=NT, dt=DT, device=dev)
    for k, (steers, xele, _) in enumerate(geos):
        xs, zs, amp = sim.scatterers(50, 1000, (-19e-3, 19e-3), (2e-3, 50e-3), 7000 + k)
        batch.rf_views[k].copy_(sim.simulate_rf_torch(steers, xele, NT, xs, zs, amp, dev,
                                                      seed=7000 + k))
    torch.cuda.synchronize()
    refs = []
    for k, (side, ntx, nrx, kind, prec) in enumerate(BATCH):
        p = geos[k][2]
        rf = batch.rf_views[k].cpu().numpy()
        if kind == "das":
            pix = np.arange(0, side * side, 11)
            tabs = oracle.synthetic_tables(p["ovectx"], p["nvectx"], p["t0tx"], p["alatx"],
                                           p["doftx"], p["ovecrx"], p["nvecrx"], p["alarx"],
                                           p["c0"], p["pnts"][pix])
            img, _ = oracle.das_tabbed_mt(rf, 0.0, DT, *tabs, kind="cubic")
            refs.append((pix, img))
        else:
            cols = np.array([0, side // 3, side // 2, side - 1])
            pix = (cols[:, None] * side + np.arange(side)[None, :]).ravel()
            tabs = oracle.synthetic_tables(p["ovectx"], p["nvectx"], p["t0tx"], p["alatx"],
                                           p["doftx"], p["ovecrx"], p["nvecrx"], p["alarx"],
                                           p["c0"], p["pnts"][pix])
            imrx = oracle.das_tabbed(rf, 0.0, DT, *tabs, kind="cubic", sumtype="tx_only")
            ref = numerics.slsc_image(imrx.reshape(nrx, cols.size, side),
                                      tabs[3].reshape(nrx, cols.size, side),
                                      list(range(1, 21)), 5, -2, 1, 1)
            refs.append((cols, ref))
    return batch, refs


@pytest.mark.parametrize("tile,rchunk", [(16, 32), (8, 16), (32, 128)])
def test_ragged_batch_requests_match_oracle(ragged, tile, rchunk):
    import torch
    batch, refs = ragged
    print(f"ragged batch of {len(BATCH)} requests at nt {NT}: tile {tile}, receive chunk "
          f"{rchunk}", flush=True)
    batch.configure(tile, rchunk)
    batch.run()
    torch.cuda.synchronize()
    for k, ((side, ntx, nrx, kind, prec), (idx, ref)) in enumerate(zip(BATCH, refs)):
        out = batch.outputs[k].cpu().numpy()
        what = f"request {k}: {side}^2 px, {ntx} tx, {nrx} rx, {kind}"
        if kind == "das":
            assert_close_to_peak(out.reshape(-1)[idx], ref, 1e-5, what)
        else:
            _compare(out[idx], ref, TOL[prec], f"{what} ({prec})")


# -------------------------------------------------------- speed of sound ---
def test_sos_map_gradient_full_size_matches_fp64_central_differences():
    """d loss / d c at 32 seeded entries of the 64 x 64 map (31 PW x 128 rx x
    4096 samples, 256 x 256 pixels): float32 kernels and their adjoints
    against central differences (h = 1e-3 m/s) of the FP64 pipeline."""
    import torch
    from oracle import diff64
    from pycbf_amd.gpu.sos import SlownessGrid, SoSProblem, focus_loss, ray_delays
    print("speed of sound: 31x128x4096, 256x256 px, 64x64 map, 32 FD entries", flush=True)
    dev = torch.device("cuda", 0)
    w = sim.workload_sos_geometry()
    geo = SlownessGrid.covering(w["xlim"], w["zlim"], w["nxm"], w["nzm"])
    f = lambda a: torch.as_tensor(np.asarray(a, np.float32), device=dev)  # noqa: E731
    ctrue = sim.sos_true_map(w["nzm"], w["nxm"], w["xlim"], w["zlim"], seed=11)
    xs, zs, amp = sim.scatterers(50, 10_000, w["xlim"], w["zlim"], 11)
    ttx, trx = ray_delays(1.0 / f(ctrue), geo, f(w["steers"]), f(w["xorig"]), f(w["xele"]),
                          f(np.stack([xs, zs], 1)), 1540.0, 64)
    rf = sim.simulate_rf_delays_torch(ttx, trx, amp, w["nt"], dev, seed=11)
    prob = SoSProblem(rf=rf, steer=f(w["steers"]), xorig=f(w["xorig"]), xele=f(w["xele"]),
                      pvec=f(w["pnts"]), gridshape=w["gridshape"], geo=geo, t0=0.0, dt=DT)
    c0map = torch.full((geo.nzm, geo.nxm), 1540.0, device=dev)
    with torch.no_grad():
        e0 = float((prob.images(c0map).sum(0) ** 2).mean())
    c = c0map.clone().requires_grad_()
    loss = focus_loss(prob.images(c).double(), e0, prob.kwin)
    (gc,) = torch.autograd.grad(loss, [c])
    gc = gc.double().cpu().numpy()

    rf64 = prob.rf.double()[None]
    t32, dt32, geo32 = float32_inputs(prob.t0, prob.dt, geo)
    y2 = diff64.natural_y2(rf64, dt32)
    st, xo, xe, pv = (t.double() for t in (prob.steer, prob.xorig, prob.xele, prob.pvec))
    nx, nz = prob.gridshape

    def L64(cm):
        with torch.no_grad():
            t64, r64 = diff64.ray_delays(s32 + (1.0 / cm - 1.0 / base), geo32, st, xo, xe, pv,
                                         1540.0, 64)
            im = diff64.das_tabbed(rf64, t64, torch.ones_like(t64), r64, torch.ones_like(r64),
                                   t32, dt32, "cubic", "tx_only", 0.0, x32=False, y2=y2)
            return float(focus_loss(im[0].reshape(-1, nx, nz), e0, prob.kwin))

    base = c0map.double()
    # the kernels take the slowness as float32(1 / c); central differences of
    # h = 1e-3 m/s lie far below float32 resolution, so the FP64 pipeline
    # expands around the kernels' base slowness
    s32 = (1.0 / c0map).double()
    rng = np.random.default_rng(5)
    h = 1e-3
    peak = float(np.abs(gc).max())
    errs = []
    for n in range(32):
        i, k = int(rng.integers(0, geo.nzm)), int(rng.integers(0, geo.nxm))
        cp, cm = base.clone(), base.clone()
        cp[i, k] += h
        cm[i, k] -= h
        fd = (L64(cp) - L64(cm)) / (2 * h)
        errs.append((abs(gc[i, k] - fd) / max(abs(fd), 1e-2 * peak), i, k, gc[i, k], fd))
        if n % 8 == 7:
            print(f"  {n + 1} of 32 entries", flush=True)
    errs.sort(reverse=True)
    print("full-size sos gradient vs FP64 central differences: worst rel errs",
          [f"{e[0]:.2e}" for e in errs[:6]], "median", f"{errs[16][0]:.2e}", flush=True)
    assert errs[0][0] <= 1e-3, errs[0]
