"""Host-side Python access to the MI355X SCGF hot path (libbcdor_amd.so).

synth    -- synthetic closed-shell systems (model space + interaction)
formats  -- BoccaDorata file formats (model space, .bcd, propagator, pivots, bcdwk)
capi     -- ctypes bindings of the C ABI in include/bcdor_amd.h
"""
from . import formats, synth  # noqa: F401
