"""pycbf_amd -- MI355X (gfx950) engine for pycbf's beamforming hot path.

``pycbf_amd.gpu`` holds the GPU beamformer classes (same API as ``pycbf.gpu``);
``pycbf_amd.base`` the shared base classes and ``BeamformerException``.
"""
from .base import BeamformerException  # noqa: F401

__version__ = "0.1.0"
