"""MI355X-native implementation of the OSCARS synchrotron-radiation hot path.

``oscars_amd.sr.sr`` is the drop-in simulation object; ``oscars_amd._lib``
holds the ctypes bindings of the C ABI declared in include/oscars_mi.h.
"""
__version__ = "0.1.0"
