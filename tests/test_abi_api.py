"""CPU tests of the boundary: libpycbf_hip.so loads and exports every entry
point include/pycbf_hip.h declares; the host-only entry points work without a
GPU; argument errors come back as status codes with a message; and the
Python classes validate like pycbf/__bf_base_classes__.py:55-219 and
pycbf/__advanced_bf_classes__.py:17-25 before touching a device."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pycbf_hip.h")
LIB = os.path.join(ROOT, "pycbf_amd", "libpycbf_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pycbf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("pycbf_das_tabbed", "pycbf_das_synthetic", "pycbf_spline_coefs", "pycbf_dmas",
                 "pycbf_synthetic_pack", "pycbf_last_error", "pycbf_das_workspace_bytes"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from pycbf_amd.gpu import _engine as E
    missing = [n for n in declared_functions() if not hasattr(E.lib, n)]
    assert not missing, f"not exported: {missing}"
    # and every symbol the Python binding types is declared in the header
    assert set(E._SIGS) <= set(declared_functions())


def test_abi_version_matches_header():
    from pycbf_amd.gpu import _engine as E
    m = re.search(r"#define PYCBF_ABI_VERSION (\d+)", open(HEADER).read())
    assert E.lib.pycbf_abi_version() == int(m.group(1)) == E.ABI_VERSION


def test_header_constants_match_binding():
    from pycbf_amd.gpu import _engine as E
    text = open(HEADER).read()
    consts = dict((k, int(v)) for k, v in re.findall(r"#define (PYCBF_\w+) (\d+)", text))
    for name, code in E.INTERP.items():
        assert consts["PYCBF_INTERP_" + name.upper()] == code
    for name, code in E.SUMTYPE.items():
        assert consts["PYCBF_SUM_" + name.upper()] == code
    for name, code in E.DMASTYPE.items():
        assert consts["PYCBF_DMAS_" + name.upper()] == code
    for name, code in E.TABS.items():
        assert consts["PYCBF_TABS_" + name.upper()] == code
    assert consts["PYCBF_TXPACK_FLOATS"] == E.TXPACK_FLOATS
    assert consts["PYCBF_RXPACK_FLOATS"] == E.RXPACK_FLOATS
    assert ct.sizeof(E.RFInfo) == 32
    assert E.RFInfo.tInfo.offset == 20 and E.RFInfo.ndim.offset == 16


def test_integration_stub_matches_header():
    """The reference-side binding in INTEGRATION.md sizes the geometry packs
    and checks the ABI version with constants of its own."""
    text = open(HEADER).read()
    consts = dict((k, int(v)) for k, v in re.findall(r"#define (PYCBF_\w+) (\d+)", text))
    stub = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"_TXPACK, _RXPACK, _ABI = (\d+), (\d+), (\d+)", stub)
    assert m, "pack constants not found in the stub"
    assert int(m.group(1)) == consts["PYCBF_TXPACK_FLOATS"]
    assert int(m.group(2)) == consts["PYCBF_RXPACK_FLOATS"]
    assert int(m.group(3)) == consts["PYCBF_ABI_VERSION"]


def test_rfinfo_layout_matches_reference_dtype():
    """struct RFInfo as the reference's numpy dtype lays it out
    (pycbf/gpu/__engines__.py:13-14)."""
    from pycbf_amd.gpu import _engine as E
    xinfo = np.dtype([("x0", np.float32), ("dx", np.float32), ("nx", np.int32)])
    rfinfo = np.dtype([("ntx", np.int64), ("nrx", np.int64), ("ndim", np.int32),
                       ("tInfo", xinfo)])
    a = np.zeros(1, rfinfo)
    a["ntx"], a["nrx"], a["ndim"] = 75, 128, 2
    a["tInfo"]["x0"], a["tInfo"]["dx"], a["tInfo"]["nx"] = 1e-6, 2.5e-8, 4096
    info = E.rfinfo(75, 128, 2, 1e-6, 2.5e-8, 4096)
    assert bytes(info) == a.tobytes()


def test_tile_order_is_a_compact_permutation():
    from pycbf_amd.gpu import _engine as E
    nx, nz = 96, 70
    x, z = np.meshgrid(np.linspace(-0.02, 0.02, nx), np.linspace(0.002, 0.05, nz), indexing="ij")
    p = np.ascontiguousarray(np.stack([x.ravel(), z.ravel()], 1), np.float32)
    perm = np.empty(p.shape[0], np.int32)
    assert E.lib.pycbf_tile_order(E.ptr(p), p.shape[0], 2, E.ptr(perm)) == 0
    assert np.array_equal(np.sort(perm), np.arange(p.shape[0]))
    # each 256-pixel tile of the order covers a small patch of the grid
    ix, iz = np.unravel_index(perm, (nx, nz))
    spans = [(np.ptp(ix[s:s + 256]) + 1) * (np.ptp(iz[s:s + 256]) + 1)
             for s in range(0, perm.size - 255, 256)]
    assert np.median(spans) <= 4 * 256


def test_host_entry_points_report_argument_errors():
    from pycbf_amd.gpu import _engine as E
    f = np.zeros(16, np.float32)
    rc = E.lib.pycbf_synthetic_pack(4, 1, E.ptr(f), E.ptr(f), E.ptr(f), E.ptr(f), E.ptr(f), 1,
                                    E.ptr(f), E.ptr(f), E.ptr(f), E.ptr(f), E.ptr(f))
    assert rc != 0
    assert b"ndim" in E.lib.pycbf_last_error()
    perm = np.empty(4, np.int32)
    assert E.lib.pycbf_tile_order(None, 4, 2, E.ptr(perm)) != 0
    assert b"NULL" in E.lib.pycbf_last_error()
    # the tab generator validates its arguments before any device work
    assert E.lib.pycbf_make_tabs_2d(7, 1, None, 1, None, None, 1, None, 1, 1540.0, 1.5, None,
                                    None, None, None, None) != 0
    assert b"mode" in E.lib.pycbf_last_error()
    assert E.lib.pycbf_make_tabs_2d(0, 1, None, 1, None, None, 1, None, 1, 1540.0, 1.5, None,
                                    None, None, None, None) != 0
    assert b"NULL" in E.lib.pycbf_last_error()
    # so does the demodulation product
    assert E.lib.pycbf_demod_apply(None, 4, 16, None, 8, None, None, 0, None) != 0
    assert b"NULL" in E.lib.pycbf_last_error()
    assert E.lib.pycbf_demod_apply(None, 4, 0, None, 8, None, None, 0, None) != 0
    assert b"n >= 1" in E.lib.pycbf_last_error()
    assert E.lib.pycbf_demod_apply(None, 0, 16, None, 8, None, None, 0, None) == 0
    # one image (512 lines x 4096 samples): a split product needs a workspace
    need = E.lib.pycbf_demod_workspace_bytes(512, 4096, 718)
    assert need >= 2 * 512 * 718 * 4 and need % (512 * 718 * 4) == 0
    assert E.lib.pycbf_demod_workspace_bytes(32768, 2048, 730) == 0  # tiles fill the GPU
    buf = np.zeros(4, np.float32)
    assert E.lib.pycbf_demod_apply(E.ptr(buf), 512, 4096, E.ptr(buf), 718, E.ptr(buf), None, 0,
                                   None) != 0
    assert b"workspace" in E.lib.pycbf_last_error()


def test_adjoint_entry_point_validates_before_device_work():
    """pycbf_das_tabbed_bwd rejects unsupported kinds, bad k-order arguments,
    oversized upsampling and short workspaces with a message, before any
    device call (host buffers stand in for device pointers: never read)."""
    from pycbf_amd.gpu import _engine as E
    info = E.rfinfo(2, 3, 0, 0.0, 2.5e-8, 64)
    buf = np.zeros(4096, np.float32)
    p = E.ptr(buf)

    def call(kind, param=1, S=None, ws=0):
        return E.lib.pycbf_das_tabbed_bwd(ct.byref(info), p, 1, p, p, p, p, None, None, 16, kind,
                                          param,
                                          S, 0.0, E.SUMTYPE["tx_and_rx"], p, p, None, None,
                                          None, None, p if ws else None, ws, None)

    assert call(99) != 0
    assert b"adjoint supports" in E.lib.pycbf_last_error()
    assert call(E.INTERP["korder_cubic"], 5, p) != 0
    assert b"even" in E.lib.pycbf_last_error()
    assert call(E.INTERP["korder_cubic"], 8, None) != 0
    assert b"S operator" in E.lib.pycbf_last_error()
    assert call(E.INTERP["nearest"], 1 << 28) != 0
    assert b"int32" in E.lib.pycbf_last_error()
    for kind, param in (("cubic", 1), ("korder_cubic", 8), ("akima", 1), ("nearest", 4)):
        assert call(E.INTERP[kind], param, p if kind == "korder_cubic" else None, ws=64) != 0
        assert b"workspace" in E.lib.pycbf_last_error()


def test_adjoint_workspace_grows_with_kind_and_upsampling():
    from pycbf_amd.gpu import _engine as E
    info = E.rfinfo(5, 16, 0, 0.0, 2.5e-8, 1024)
    q = lambda kind, param=1: E.lib.pycbf_das_bwd_workspace_bytes(  # noqa: E731
        ct.byref(info), 2, 4096, E.INTERP[kind], param)
    nearest, up4 = q("nearest"), q("nearest", 4)
    # the upsampled trace and its cotangents: 2 x nch x (4 (nt-1) + 1) floats
    assert up4 - nearest >= 2 * 2 * 5 * 16 * (4 * 1023 + 1) * 4
    # Hermite kinds: a float4 table and four cotangent rows per channel
    assert q("korder_cubic", 8) >= nearest + 2 * 5 * 16 * 1024 * (16 + 4 * 4)
    assert q("akima") == q("makima") == q("korder_cubic", 8)


def test_workspace_query_is_host_only_and_grows_with_frames():
    from pycbf_amd.gpu import _engine as E
    info = E.rfinfo(5, 128, 0, 0.0, 2.5e-8, 2048)
    one = E.lib.pycbf_das_workspace_bytes(ct.byref(info), 1, 32768, E.INTERP["cubic"], 1, 0)
    four = E.lib.pycbf_das_workspace_bytes(ct.byref(info), 4, 32768, E.INTERP["cubic"], 1, 0)
    # the cubic table alone is nfr * ntx * nrx * (nt + 1) * 8 bytes
    assert one >= 5 * 128 * 2049 * 8
    assert four >= 4 * 5 * 128 * 2049 * 8 and four > one


# ---------------------------------------------------------------- validation
def _tabs(ntx=2, nrx=3, np_=5):
    return dict(tautx=np.zeros((ntx, np_)), taurx=np.zeros((nrx, np_)),
                apodtx=np.ones((ntx, np_)), apodrx=np.ones((nrx, np_)))


def _synth(ntx=2, nrx=3, np_=5, ndim=2):
    return dict(ovectx=np.zeros((ntx, ndim)), nvectx=np.zeros((ntx, ndim)), doftx=np.zeros(ntx),
                alatx=np.zeros(ntx), t0tx=np.zeros(ntx), ovecrx=np.zeros((nrx, ndim)),
                nvecrx=np.zeros((nrx, ndim)), alarx=np.zeros(nrx), c0=1540.0,
                pnts=np.zeros((np_, ndim)))


TIME = dict(t0=0.0, dt=2.5e-8, nt=64)


@pytest.mark.parametrize("interp,msg", [
    ({"kind": "spline"}, "interp"), ({}, "interp"),
    ({"kind": "korder_cubic", "k": 5}, "even"), ({"kind": "korder_cubic", "k": 2}, "even"),
    ({"kind": "nearest", "usf": 0}, "usf"), ({"kind": "linear", "usf": 2}, "upsampling"),
])
def test_interp_validation(interp, msg):
    from pycbf_amd.base import BeamformerException
    from pycbf_amd.gpu import TabbedBeamformer
    with pytest.raises(BeamformerException, match=msg):
        TabbedBeamformer(**_tabs(), **TIME, interp=interp)


def test_interp_defaults():
    from pycbf_amd.base import Beamformer
    b = Beamformer(interp={"kind": "korder_cubic"})
    assert b.interp["k"] == 8
    b = Beamformer(interp={"kind": "nearest"})
    assert b.interp["usf"] == 1
    assert Beamformer().thresh == 1e-2 and Beamformer().sumtype == "tx_and_rx"


def test_sumtype_validation():
    from pycbf_amd.base import BeamformerException
    from pycbf_amd.gpu import TabbedBeamformer
    with pytest.raises(BeamformerException, match="sumtype"):
        TabbedBeamformer(**_tabs(), **TIME, sumtype="both")


@pytest.mark.parametrize("field,shape,msg", [
    ("tautx", (2, 5, 1), "2D"), ("apodtx", (3, 5), "dim 0"), ("apodrx", (2, 5), "dim 0"),
    ("taurx", (3, 4), "dim 1"), ("apodrx", (3, 4), "dim 1"),
])
def test_tabbed_shape_validation(field, shape, msg):
    from pycbf_amd.base import BeamformerException
    from pycbf_amd.gpu import TabbedBeamformer
    kw = _tabs()
    kw[field] = np.zeros(shape)
    with pytest.raises(BeamformerException, match=msg):
        TabbedBeamformer(**kw, **TIME)


@pytest.mark.parametrize("field,value,msg", [
    ("doftx", np.zeros((2, 1)), "1D"), ("ovecrx", np.zeros(3), "2D"),
    ("pnts", np.zeros(5), "Np by ndim"), ("t0tx", np.zeros(3), "tx variables"),
    ("alarx", np.zeros(2), "rx variables"), ("nvecrx", np.zeros((3, 3)), "dimensions"),
])
def test_synthetic_shape_validation(field, value, msg):
    from pycbf_amd.base import BeamformerException
    from pycbf_amd.gpu import SyntheticBeamformer
    kw = _synth()
    kw[field] = value
    with pytest.raises(BeamformerException, match=msg):
        SyntheticBeamformer(**kw, **TIME)


def test_synthetic_rejects_4d_points():
    from pycbf_amd.base import BeamformerException
    from pycbf_amd.gpu import SyntheticBeamformer
    with pytest.raises(BeamformerException, match="2D or 3D"):
        SyntheticBeamformer(**_synth(ndim=4), **TIME)


def test_time_axis_validation():
    from pycbf_amd.base import BeamformerException
    from pycbf_amd.gpu import TabbedBeamformer
    with pytest.raises(BeamformerException, match="dt"):
        TabbedBeamformer(**_tabs(), t0=0.0, dt=0.0, nt=64)
    with pytest.raises(BeamformerException, match="nt"):
        TabbedBeamformer(**_tabs(), t0=0.0, dt=1e-8, nt=0)


@pytest.mark.parametrize("kw,msg", [
    ({"lagaxis": "z"}, "lagaxis"), ({"dmastype": "cubic"}, "dmastype"),
    ({"lags": np.array([])}, "empty"), ({"lags": np.array([1.5])}, "integers"),
])
def test_dmas_validation(kw, msg):
    from pycbf_amd.base import BeamformerException
    from pycbf_amd.gpu import TabbedDMASBeamformer
    args = dict(lags=np.arange(1, 4))
    args.update(kw)
    with pytest.raises(BeamformerException, match=msg):
        TabbedDMASBeamformer(**_tabs(), **TIME, **args)


def test_no_device_raises_instead_of_falling_back():
    """Without a GPU the product path fails loudly (no CPU fallback)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from pycbf_amd.base import BeamformerException
    from pycbf_amd.gpu import TabbedBeamformer
    with pytest.raises(BeamformerException, match="device"):
        TabbedBeamformer(**_tabs(), **TIME)
