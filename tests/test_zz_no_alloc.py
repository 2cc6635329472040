"""pycbf_das_synthetic_noalloc (library-owned scratch, no allocation) against
pycbf_das_synthetic (caller workspace), through the C-ABI: two calls in a
row on one GPU, the second on a smaller problem, and a problem beyond the
scratch array refused with a message.  Both entry points run the same
kernels with the same launch plan: outputs of the deterministic launches (no
atomics) are bit for bit the same; the one-frame image summed over both axes,
whose partial images meet in float atomics, is within 1e-5 of its peak."""
import ctypes as ct

import numpy as np
import pytest

from pycbf_amd import sim
from pycbf_amd.gpu import _engine as E
from tests.helpers import assert_close_to_peak

pytestmark = pytest.mark.gpu


def _problem(ntx, nrx, nt, nx, nz, nfr, seed):
    import torch
    dev = torch.device("cuda", 0)
    xele = sim.linear_array(nrx)
    steers = np.radians(np.linspace(-15, 15, ntx))
    _, _, pnts = sim.grid(nx, nz, (-19e-3, 19e-3), (5e-3, 45e-3))
    p = sim.pw_synthetic(steers, xele, pnts)
    txpack = np.zeros((ntx, E.TXPACK_FLOATS), np.float32)
    rxpack = np.zeros((nrx, E.RXPACK_FLOATS), np.float32)
    E.check(E.lib.pycbf_synthetic_pack(2, ntx, E.ptr(p["ovectx"]), E.ptr(p["nvectx"]),
                                       E.ptr(p["t0tx"]), E.ptr(p["alatx"]), E.ptr(p["doftx"]),
                                       nrx, E.ptr(p["ovecrx"]), E.ptr(p["nvecrx"]),
                                       E.ptr(p["alarx"]), E.ptr(txpack), E.ptr(rxpack)), "pack")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    rf = torch.randn((nfr, ntx, nrx, nt), generator=g, device=dev)
    t = lambda a: torch.as_tensor(a, device=dev)  # noqa: E731
    return dict(info=E.rfinfo(ntx, nrx, 2, 0.0, 1.0 / sim.FS, nt), rf=rf, nfr=nfr,
                txpack=t(txpack), rxpack=t(rxpack), pvec=t(p["pnts"]), np=nx * nz, dev=dev)


def _both(P, kind, param, sumflag, shape):
    import torch
    dev = P["dev"]
    s = E.stream_of(dev)
    args = (ct.byref(P["info"]), E.ptr(P["rf"]), P["nfr"], E.ptr(P["txpack"]),
            E.ptr(P["rxpack"]), 1540.0, E.ptr(P["pvec"]), None, P["np"], E.INTERP[kind], param,
            None, 1e-2, E.SUMTYPE[sumflag])
    ref = torch.full(shape, np.nan, dtype=torch.float32, device=dev)
    wsb = E.lib.pycbf_das_workspace_bytes(ct.byref(P["info"]), P["nfr"], P["np"],
                                          E.INTERP[kind], param, 1)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    E.check(E.lib.pycbf_das_synthetic(*args, E.ptr(ref), E.ptr(ws), wsb, s), "workspace call")
    got = torch.full(shape, np.nan, dtype=torch.float32, device=dev)
    E.check(E.lib.pycbf_das_synthetic_noalloc(*args, E.ptr(got), s), "no-alloc call")
    torch.cuda.synchronize()
    return got.cpu().numpy(), ref.cpu().numpy()


def test_noalloc_matches_workspace_entry_point_twice():
    print("no-alloc: 31x128x4096 cubic, 256x256 px, 1 frame", flush=True)
    big = _problem(31, 128, 4096, 256, 256, 1, seed=31)
    got, ref = _both(big, "cubic", 1, "tx_and_rx", (1, 256 * 256))
    assert np.all(np.isfinite(ref)) and np.abs(ref).max() > 0
    # one frame summed over both axes adds partial images with float atomics:
    # the same entry point's calls agree to float32 rounding, not bit for bit
    assert_close_to_peak(got, ref, 1e-5, "cubic tx_and_rx, 1 frame")
    print("no-alloc: 5x64x2048 linear, 96x128 px, 2 frames, sumtype tx_only", flush=True)
    small = _problem(5, 64, 2048, 96, 128, 2, seed=5)
    got, ref = _both(small, "linear", 1, "tx_only", (2, 64, 96 * 128))
    assert np.all(np.isfinite(ref)) and np.abs(ref).max() > 0
    assert np.array_equal(got, ref), f"{int((got != ref).sum())} samples differ"


def test_noalloc_refuses_a_problem_beyond_its_scratch():
    import torch
    print("no-alloc: 75x128x4096 cubic is refused (table beyond the scratch)", flush=True)
    P = _problem(75, 128, 4096, 64, 64, 1, seed=75)
    out = torch.zeros((1, 64 * 64), dtype=torch.float32, device=P["dev"])
    rc = E.lib.pycbf_das_synthetic_noalloc(
        ct.byref(P["info"]), E.ptr(P["rf"]), 1, E.ptr(P["txpack"]), E.ptr(P["rxpack"]), 1540.0,
        E.ptr(P["pvec"]), None, P["np"], E.INTERP["cubic"], 1, None, 1e-2, E.SUMTYPE["tx_and_rx"],
        E.ptr(out), E.stream_of(P["dev"]))
    assert rc == 1001
    assert "scratch of" in E.last_error() and "needed" in E.last_error()
