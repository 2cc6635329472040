"""Ragged batches (pycbf_batch_*): every request of one persistent launch
against the single-request path, every autotuner configuration, and graph
replays against the eager run.

Bars: DAS images within 1e-5 of the single-request image's peak (float32
sums in another order); SLSC images (FP32) within 1e-4 absolute of SLSC on the
single-request per-receive images, NaN in the same places."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(seed=0, nt=768):
    from pycbf_amd import sim
    from pycbf_amd.gpu.batch import BatchRequest
    rng = np.random.default_rng(seed)
    shapes = [(16, 24, 1, 8, "das"), (40, 33, 3, 16, "das"), (24, 24, 5, 24, "slsc"),
              (9, 70, 2, 12, "das"), (32, 16, 4, 40, "slsc"), (48, 40, 7, 20, "das"),
              # 136 receive elements: a ragged second chunk at every receive chunk
              (12, 10, 2, 136, "das")]
    reqs = []
    for k, (nx, nz, ntx, nrx, kind) in enumerate(shapes):
        xele = sim.linear_array(nrx)
        steers = np.radians(np.linspace(-10, 10, ntx)) if ntx > 1 else np.zeros(1)
        lim = float(xele[-1]) * 1.2 + 1e-3
        _, _, pnts = sim.grid(nx, nz, (-lim, lim), (2e-3, 14e-3))
        params = sim.pw_synthetic(steers, xele, pnts)
        xs, zs, amp = sim.scatterers(6, 60, (-lim, lim), (3e-3, 13e-3), seed + k)
        rf = sim.simulate_rf(steers, xele, nt, xs, zs, amp, seed=seed + k).astype(np.float32)
        reqs.append((BatchRequest(params=params, gridshape=(nx, nz), kind=kind,
                                  lags=tuple(range(1, 6)), kwin=4, dsfz=2, precision="fp32"),
                     rf))
    return reqs


def _single(req, rf, nt):
    """The same request through SyntheticBeamformer (+ slsc)."""
    import torch
    from pycbf_amd import sim
    from pycbf_amd.gpu import SyntheticBeamformer, slsc
    nx, nz = req.gridshape
    if req.kind == "das":
        bf = SyntheticBeamformer(**req.params, t0=0.0, dt=1.0 / sim.FS, nt=nt,
                                 interp={"kind": "cubic"}, sumtype="tx_and_rx")
        return bf(rf).reshape(nx, nz)
    bf = SyntheticBeamformer(**req.params, t0=0.0, dt=1.0 / sim.FS, nt=nt,
                             interp={"kind": "cubic"}, sumtype="tx_only")
    imrx = bf(rf, out_as_numpy=False)
    apod = bf.tables()[3]
    nrx = imrx.shape[0]
    return slsc(imrx.view(nrx, nx, nz), req.lags, req.kwin, -(req.kwin // 2),
                apod.view(nrx, nx, nz), req.dsfx, req.dsfz, "fp32").cpu().numpy()


def _check(batch, reqs, refs):
    for r, ((req, _), ref) in enumerate(zip(reqs, refs)):
        got = batch.outputs[r].cpu().numpy()
        assert got.shape == ref.shape, r
        if req.kind == "das":
            peak = np.max(np.abs(ref))
            assert np.max(np.abs(got - ref)) <= 1e-5 * peak, r
        else:
            assert np.array_equal(np.isnan(got), np.isnan(ref)), r
            fin = ~np.isnan(ref)
            assert np.max(np.abs(got[fin] - ref[fin])) <= 1e-4, r


@pytest.fixture(scope="module")
def setup():
    import torch
    from pycbf_amd import sim
    from pycbf_amd.gpu.batch import RaggedBatch
    nt = 768
    reqs = _requests(nt=nt)
    refs = [_single(q, rf, nt) for q, rf in reqs]
    batch = RaggedBatch([q for q, _ in reqs], nt=nt, dt=1.0 / sim.FS)
    for v, (_, rf) in zip(batch.rf_views, reqs):
        v.copy_(torch.as_tensor(rf))
    return batch, reqs, refs


def test_batch_matches_single_requests(setup):
    import torch
    batch, reqs, refs = setup
    print("ragged batch of 7 requests, default configuration", flush=True)
    batch.run()
    torch.cuda.synchronize()
    _check(batch, reqs, refs)
    lat = batch.latencies_ms()
    assert np.all(lat > 0) and np.all(np.isfinite(lat))


@pytest.mark.parametrize("tile", [8, 16, 32])
@pytest.mark.parametrize("rchunk", [16, 32, 64, 128])
def test_batch_every_configuration(setup, tile, rchunk):
    import torch
    batch, reqs, refs = setup
    print(f"ragged batch tile {tile} rchunk {rchunk}", flush=True)
    batch.configure(tile, rchunk)
    batch.run()
    torch.cuda.synchronize()
    _check(batch, reqs, refs)


def test_autotune_then_graph_replays(setup):
    import torch
    batch, reqs, refs = setup
    tile, rchunk, times = batch.autotune(reps=1)
    assert (tile, rchunk) in times and times[(tile, rchunk)] == min(times.values())
    batch.run()
    torch.cuda.synchronize()
    eager = [o.clone() for o in batch.outputs]
    batch.capture()
    for k in range(3):
        for o in batch.outputs:      # the program's own eager work between replays
            o.fill_(-7.0) if o.dtype == torch.float32 else None
        torch.cuda.synchronize()
        batch.run()
        torch.cuda.synchronize()
        for r, (o, e) in enumerate(zip(batch.outputs, eager)):
            assert torch.equal(torch.isnan(o), torch.isnan(e)), (k, r)
            ok = ~torch.isnan(e)
            assert float((o[ok] - e[ok]).abs().max()) <= 1e-5 * float(e[ok].abs().max() + 1e-30), (k, r)
    _check(batch, reqs, refs)
