"""`beampower.core` of the reference (beampower/core.py): the library
table and loader, from beampower_amd."""
from beampower_amd.core import LIBRARIES, load_library  # noqa: F401

__all__ = ["LIBRARIES", "load_library"]
