"""beampower -- the reference package's import name (beampower v1.0.4,
beampower/__init__.py:9-14), served by this repository's gfx950 engine.

`import beampower as bp; bp.beamform(..., device="gpu")` runs the
MI355X library; the module layout is the reference's: `beampower.beamform`,
`beampower.load_library`, `beampower.beampower.beamform`,
`beampower.beampower.prestack_traces` and `beampower.core`.  The code lives
in the package `beampower_amd`; these modules only re-export it.
"""
from beampower_amd import __version__  # noqa: F401
from beampower_amd.beampower import beamform
from beampower_amd.core import load_library

from . import beampower, core  # noqa: F401,E402  (the reference's submodules)

__all__ = ["load_library", "beamform"]
