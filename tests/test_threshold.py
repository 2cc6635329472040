"""The threshold that reproduces pycbf's GPU weight rule.

pycbf's GPU kernels keep a sample when ``abs(apodtx * apodrx) > 0.01``: a
float compared with the double literal 0.01 (pycbf/gpu/__engines__.cu:8,564,
678,762,851).  The library keeps a sample when ``|w| >= thresh`` (pycbf's CPU
rule, pycbf/cpu/pycbf.c:201,267).  ``PYCBF_GPU_REFERENCE_THRESH``, the
smallest float above 0.01, makes the two rules agree on every float; the
integration stub passes it.  Checked here on every float between 0.005 and
0.02 of both signs, plus the special values."""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pycbf_hip.h")


def _header_thresh():
    m = re.search(r"#define PYCBF_GPU_REFERENCE_THRESH ([0-9.eE+-]+)f", open(HEADER).read())
    return np.float32(float(m.group(1)))


def test_header_and_binding_constant_are_the_float_above_one_hundredth():
    from pycbf_amd.gpu import _engine as E
    t = _header_thresh()
    assert t == np.float32(E.GPU_REFERENCE_THRESH)
    assert t == np.nextafter(np.float32(0.01), np.float32(1.0))
    assert float(t) > 0.01 > float(np.nextafter(t, np.float32(0.0)))


def test_rule_equivalence_on_every_float_near_one_hundredth():
    t = _header_thresh()
    lo, hi = np.float32(0.005).view(np.int32), np.float32(0.02).view(np.int32)
    w = np.arange(lo, hi + 1, dtype=np.int32).view(np.float32)
    w = np.concatenate([w, -w, np.float32([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, 1e-30])])
    gpu_rule = np.abs(w).astype(np.float64) > 0.01           # NaN compares false
    lib_rule = np.abs(w) >= t
    assert np.array_equal(gpu_rule, lib_rule)
    # the CPU default 1e-2 differs exactly at float32(0.01)
    cpu_rule = np.abs(w) >= np.float32(1e-2)
    diff = w[cpu_rule != gpu_rule]
    assert set(np.abs(diff).tolist()) == {float(np.float32(0.01))}
