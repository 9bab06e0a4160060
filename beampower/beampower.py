"""`beampower.beampower` of the reference (beampower/beampower.py): the
beamforming call and the channel prestack, from beampower_amd."""
from beampower_amd.beampower import beamform, prestack_traces  # noqa: F401

__all__ = ["beamform", "prestack_traces"]
