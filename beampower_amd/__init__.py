"""beampower_amd -- gfx950 (MI355X) beamforming engine.

Keeps the public call of the reference package `beampower` v1.0.4
(beampower/__init__.py:9-14): `beamform`, `load_library` and the
`beampower.beamform` / `beampower.prestack_traces` module layout.
"""

__all__ = ["load_library", "beamform"]

from .core import load_library
from .beampower import beamform, prestack_traces  # noqa: F401

__version__ = "1.0.4+mi355x.0.1.0"
