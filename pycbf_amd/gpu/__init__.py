"""GPU beamformers with pycbf's ``pycbf.gpu`` API, running hand-written gfx950
kernels through libpycbf_hip.so (pycbf/gpu/__init__.py:1-3)."""
from .core import SyntheticBeamformer, TabbedBeamformer, make_S_by_k  # noqa: F401
from .advanced import SyntheticDMASBeamformer, TabbedDMASBeamformer  # noqa: F401
from .advanced import SyntheticSLSCBeamformer, TabbedSLSCBeamformer, slsc, slsc_selection  # noqa: F401,E501
